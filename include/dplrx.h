/*
 * dplrx.h -- device-resident extension API of the MI355X PLR library.
 *
 * The drop-in calls of dplr.h take host memory and re-upload the matrix on
 * every call, because the caller may change the tree between calls.  These
 * entry points have no counterpart in the reference: they keep a matrix in
 * HBM as a "plan" (level-wise arrays, see dplr_hip.h) and take device
 * pointers, so a caller can apply and solve many right-hand sides without
 * host traffic.  Every function returns 0 on success and a negative code on
 * failure, with a message from dplrx_last_error().  `stream` is a
 * hipStream_t (NULL = the null stream).  Calls on one plan must not overlap
 * in time: the plan's workspace is shared.
 */
#ifndef DPLR_AMD_DPLRX_H
#define DPLR_AMD_DPLRX_H

#include <stddef.h>

#include "dplr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dplrh_plan dplrx_plan;

/* Upload the m x m uniform PLR tree A. */
int dplrx_plan_create(int m, DPlrMtx *A, dplrx_plan **plan);
void dplrx_plan_destroy(dplrx_plan *plan);
/* Order, depth, leaf order and the stored rank of each level (ranks must
 * hold depth entries; pass NULL to skip). */
int dplrx_plan_info(const dplrx_plan *plan, int *m, int *nlvl, int *leaf, int *ranks);
/* Bytes of the FP64 level arrays (U, V, sig, leaves) held in HBM. */
size_t dplrx_plan_bytes(const dplrx_plan *plan);

/* y := alpha*op(A)*x + beta*y with device vectors of length m.  trans as in
 * dplrgepmv, including the alpha == 0 quirk. */
int dplrx_gepmv(dplrx_plan *plan, char trans, double alpha, const double *dx,
                double beta, double *dy, void *stream);

/* dplrx_gepmv without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * kernels' partial sums, coefficients and arrival counters live in a static
 * device array of DPLRX_SCRATCH_BYTES on each GPU, held until the process
 * ends and shared by every call on that GPU: calls that overlap in time on
 * one GPU race.  Returns -7 (message in dplrx_last_error) when the plan
 * needs more scratch than that: with leaves of a multiple of 64 rows it
 * needs 8 * (N/128 * sum of R_d over the levels with segments longer than
 * 128 rows + sum of 2^(d+1) R_d) bytes, 1.5 MB for N = 65536 with rank 32
 * on 10 levels, 3.4 MB for N = 131072 with rank 32 on 11.  The plan must
 * have been created on the current device. */
#define DPLRX_SCRATCH_BYTES ((size_t)4 << 20)
int dplrx_gepmv_noalloc(dplrx_plan *plan, char trans, double alpha, const double *dx,
                        double beta, double *dy, void *stream);

/* C := alpha*op(A)*B + beta*C with device column-major m x n blocks
 * (ldb, ldc >= m). */
int dplrx_gepdmm(dplrx_plan *plan, char transa, int n, double alpha, const double *dB,
                 int ldb, double beta, double *dC, int ldc, void *stream);

/* Factor A in HBM for solves: FP64 leaf LU and inverses, Dinv-applied U
 * factors and one inverted capacitance matrix per node.  *info as in
 * dplrgeinv (> 0: singular leaf or capacitance).  Returns after the
 * factorization is complete; part of it runs on a second stream of the
 * plan (default priority), ordered after `stream` and joined back to it by
 * events, so the call is not for graph capture. */
int dplrx_factor(dplrx_plan *plan, int *info, void *stream);

/* X := op(A)^{-1} B for n device columns (ldb, ldx >= m), after
 * dplrx_factor.  X must not alias B. */
int dplrx_solve(dplrx_plan *plan, char trans, int n, const double *dB, int ldb, double *dX,
                int ldx, void *stream);

/* Mixed-precision factors for refined solves: f16 (type 1) or bf16
 * (type 2) U/V factors and leaf inverses with FP32 accumulation on MFMA,
 * FP64 leaf LU and capacitance LU, FP64 capacitance inverses on the top two
 * levels.  Type 3 stores each factor entry as a pair of f16 values (hi + lo
 * of the entry times a power of two per level, taken from the FP64
 * factorization, which runs first when the plan has none): three f16 MFMAs
 * per product carry about 22 bits, so a refined solve to 1e-10 takes two
 * solves instead of three (f16) or four (bf16); it needs leaves of 64 to
 * 512 rows and ranks of at most 64.  Needs the leaf order to be a multiple
 * of 64. */
int dplrx_factor_mixed(dplrx_plan *plan, int type, int *info, void *stream);

/* X := A^{-1} B by `iters` mixed-precision solves with FP64 residuals
 * r = B - A X between them (iterative refinement).  res (host, n entries,
 * optional) receives max|B - A X| per column after the last solve.
 * Speed: every refinement step costs one FP64 product, and on MI355X the
 * FP64 dplrx_solve costs about one such product, so to 1e-10 dplrx_solve is
 * the faster call at every shape measured (N = 65536 and 131072, ranks 32
 * and 64, 64 columns: f16 pairs 0.51x, f16 0.33-0.41x, bf16 0.24-0.30x of
 * its throughput); one 16-bit solve (iters = 1, ~1e-4) is the fast one.
 * Fast paths: B 16-byte aligned with an even ldb is read in place, and the
 * residuals come out of the fused FP64 product when X is 16-byte aligned
 * with an even ldx too; other layouts take a copy of B and separate passes
 * (same results). */
int dplrx_solve_refined(dplrx_plan *plan, int n, const double *dB, int ldb, double *dX,
                        int ldx, int iters, double *res, void *stream);

/* X := A^{-1} B refined until the estimated normwise relative error
 * |X - A^{-1}B|_F / |A^{-1}B|_F is at most tol, with at most maxit
 * (2 <= maxit <= 16) low-precision solves.  After solve k >= 2 the estimate is
 * 1.25 * delta_k * min(1, delta_k / delta_{k-1}) with delta_k =
 * |D_k|_F / |X_k|_F the relative size of the k-th correction D_k
 * (delta_0 = 1): a correction measures the error of the previous iterate,
 * and its ratio to the one before the contraction of the refinement (the
 * factor 1.25 covers the contraction's growth from step to step).  Columns are taken in chunks of
 * 128, each with its own estimate.  *iters (optional) receives the most
 * solves a chunk made, *err (optional) the largest estimate (compare it
 * with tol: 0 is returned also when maxit solves did not reach it).  Reads
 * the estimate back (a synchronisation of `stream`) after each solve from
 * the second on, or, when the previous call on the plan had the same tol,
 * first after as many solves as that call made.
 * Fast paths: B 16-byte aligned with an even ldb is read in place and the
 * residuals come out of the fused FP64 product when X is also 16-byte
 * aligned with an even ldx; other layouts go through a copy of B. */
int dplrx_solve_refined_tol(dplrx_plan *plan, int n, const double *dB, int ldb, double *dX,
                            int ldx, double tol, int maxit, int *iters, double *err,
                            void *stream);

/* The mixed-precision apply and solve of one step in one call:
 * Ya := A Xa (FP64) and X := A^{-1} B refined to tol as
 * dplrx_solve_refined_tol (same results, estimate and arguments).  No
 * counterpart in the reference, whose apply and inverse are separate calls
 * (src/dplrpdmm.c:454, src/dplrinv.c:93).  The apply rides on the first
 * residual product, r = B - A X_1, when that product takes the fused path
 * (n <= 64 per chunk, B and Xa 16-byte aligned with even leading
 * dimensions, leaf orders 64 * 2^k, N a multiple of 256): one pass over the
 * factors and one launch of each kernel serve both (measured: a 763-756 us
 * step against 768-771 us as two calls at N=65536, 64 columns; the
 * 128-column product is paced by its FP64 MFMA work, so the shared reads
 * save little).  Otherwise the apply is a separate FP64 product.  Ya must
 * not alias Xa, B or X. */
int dplrx_apply_solve_refined_tol(dplrx_plan *plan, int n, const double *dXa, int ldxa,
                                  double *dYa, int ldya, const double *dB, int ldb, double *dX,
                                  int ldx, double tol, int maxit, int *iters, double *err,
                                  void *stream);

/* ---- single precision --------------------------------------------------- */

/* Single-precision copies of the factors for FP32 products and solves:
 * type 1 stores U, V, the leaves and (once factored) the Dinv-applied U and
 * the leaf inverses in FP32, type 2 the same matrices in BF16; sig and the
 * capacitance inverses are FP32 in both.  The pivoted leaf LU and every
 * capacitance LU run in FP64 (dplrx_factor, run first when the plan has no
 * factorization); the products and solves then run in FP32 on FP32 MFMA
 * with FP32 vectors.  *info as dplrx_factor. */
int dplrx_factor_single(dplrx_plan *plan, int type, int *info, void *stream);
/* C := alpha*op(A)*B + beta*C with FP32 device blocks, after
 * dplrx_factor_single of `type`.  ldb, ldc >= m and multiples of 4, the
 * blocks 16-byte aligned. */
int dplrx_gepdmm_single(dplrx_plan *plan, int type, char trans, int n, float alpha,
                        const float *dB, int ldb, float beta, float *dC, int ldc, void *stream);
/* X := op(A)^{-1} B with FP32 device blocks as above (X must not alias B). */
int dplrx_solve_single(dplrx_plan *plan, int type, char trans, int n, const float *dB, int ldb,
                       float *dX, int ldx, void *stream);

/* ---- adjoints ----------------------------------------------------------- */

/* Gradient of <G, A> with respect to A's factors for G = alpha * Ga Gb^T
 * (Ga, Gb: device m x n blocks, lda, ldb >= m), written in the plan's
 * layouts: dU and dV like the level-wise U and V (column-major m x sum_d
 * R_d), dS like sig (sum_d 2^(d+1) R_d), dD like the leaves (m x leaf).
 * For y = A^{-1} b and a loss with gradient ybar: Ga = A^{-T} ybar (a
 * transposed dplrx_solve, which is also the gradient with respect to b),
 * Gb = y, alpha = -1.  For y = A x: Ga = ybar, Gb = x, alpha = 1. */
int dplrx_grad_outer(dplrx_plan *plan, int n, double alpha, const double *dGa, int lda,
                     const double *dGb, int ldb, double *dU, double *dV, double *dS, double *dD,
                     void *stream);

/* dplrx_grad_outer in FP32 on the plan's FP32 factors (dplrx_factor_single
 * type 1), FP32 blocks and outputs. */
int dplrx_grad_outer_f32(dplrx_plan *plan, int n, float alpha, const float *dGa, int lda,
                         const float *dGb, int ldb, float *dU, float *dV, float *dS, float *dD,
                         void *stream);

/* ---- batches of independent matrices ----------------------------------- */

/* A batch over `count` existing plans (not owned; they must outlive the
 * batch).  Every plan needs leaves of a multiple of 64 rows.  One call
 * runs every matrix at once: one kernel launch per phase covers all of
 * them, and the sequence is replayed from a HIP graph on later calls. */
typedef struct dplrh_batch dplrx_batch;
int dplrx_batch_create(int count, dplrx_plan *const *plans, dplrx_batch **batch);
void dplrx_batch_destroy(dplrx_batch *batch);
/* Y[i] := A_i X[i] for nrhs device columns per matrix (host arrays of
 * device pointers and leading dimensions; leading dimensions even). */
int dplrx_batch_gepdmm(dplrx_batch *batch, int nrhs, const double *const *dX, const int *ldx,
                       double *const *dY, const int *ldy, void *stream);
/* X[i] := A_i^{-1} B[i], every plan factored with dplrx_factor. */
int dplrx_batch_solve(dplrx_batch *batch, int nrhs, const double *const *dB, const int *ldb,
                      double *const *dX, const int *ldx, void *stream);
/* Autotuner: times the batch's apply + solve (nrhs columns, the caller's
 * buffers X and Y, `reps` repetitions after a warm-up) for each candidate
 * tile shape (rows per reduction task: 128, 256, 512, 1024, 2048, 4096) and
 * keeps the fastest for later calls.  *tile receives the choice; ms[0..5]
 * (optional) the time per repetition of each candidate.  No reference
 * counterpart: the reference has no batched path. */
int dplrx_batch_tune(dplrx_batch *batch, int nrhs, const double *const *dX, const int *ldx,
                     double *const *dY, const int *ldy, int reps, int *tile, double *ms,
                     void *stream);
/* Sets the tile shape (one of the tuner's candidates) of every precision
 * without timing. */
int dplrx_batch_set_tile(dplrx_batch *batch, int tile);
/* Any precision and direction.  prec 0: FP64 (double vectors, plans
 * factored with dplrx_factor for solves); 1: FP32, 2: BF16 factor matrices
 * with FP32 arithmetic (float vectors, leading dimensions multiples of 4,
 * blocks 16-byte aligned, plans prepared with dplrx_factor_single of type
 * prec).  op 0: Y[i] := alpha*op(A_i)*X[i] + beta*Y[i]; op 1:
 * Y[i] := op(A_i)^{-1} X[i].  trans 'N' or 'T'. */
int dplrx_batch_op(dplrx_batch *batch, int prec, int op, char trans, int nrhs,
                   const void *const *dX, const int *ldx, void *const *dY, const int *ldy,
                   double alpha, double beta, void *stream);
/* X[i] := A_i^{-1} B[i] by `iters` solves with the BF16 factors refined with
 * FP32 residuals B - A X (FP32 factors): every plan prepared with
 * dplrx_factor_single of types 1 and 2.  FP32 blocks as dplrx_batch_op. */
int dplrx_batch_solve_refined(dplrx_batch *batch, int nrhs, const float *const *dB,
                              const int *ldb, float *const *dX, const int *ldx, int iters,
                              void *stream);
/* dplrx_batch_tune for precision prec (vectors as dplrx_batch_op). */
int dplrx_batch_tune_prec(dplrx_batch *batch, int prec, int nrhs, const void *const *dX,
                          const int *ldx, void *const *dY, const int *ldy, int reps, int *tile,
                          double *ms, void *stream);

const char *dplrx_last_error(void);

/* ---- PLR files (no reference counterpart: the reference keeps trees in
 * memory only) ----------------------------------------------------------- */

/* Writes the m x m tree A to `path`: header "DPLRTREE", version 1, the
 * order, then every node in pre-order (type, size; children nw ne se sw;
 * leaves column-major; low-rank blocks rank, u, v, sig -- valid columns
 * only), little-endian.  0 or a negative code. */
int dplrx_save(const char *path, int m, DPlrMtx *A);
/* Replaces A (dplrini'd) with the tree stored in `path`; *m (optional) gets
 * its order.  A file that is truncated, has trailing bytes or sizes that do
 * not halve down the tree is refused (negative code, A left empty). */
int dplrx_load(const char *path, DPlrMtx *A, int *m);

/* ---- synthetic inputs (deterministic, counter-based RNG) --------------- */

/* Random uniform PLR tree of order m = leaf * 2^nlvl with every off-diagonal
 * block U*diag(s)*V^T, U and V orthonormal (Householder QR of N(0,1)
 * matrices), s_j = 10^(-j/4), and leaves N(0,1)/sqrt(leaf) + shift*I.
 * A must be dplrini'd; it is reshaped. */
int dplrx_gen_random(int m, int nlvl, int rank, double shift, unsigned long long seed,
                     DPlrMtx *A);
/* n samples of N(0,1) from stream `stream` of `seed`. */
int dplrx_gen_normal(size_t n, unsigned long long seed, unsigned int stream, double *out);

#ifdef __cplusplus
}
#endif

#endif
