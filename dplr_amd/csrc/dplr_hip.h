/*
 * Internal interface between the C host layer (csrc/host) and the HIP layer
 * (csrc/hip).  Plain C; no HIP or torch types.
 *
 * Level-wise layout.  A uniform PLR tree of order N = b * 2^L has L levels
 * of low-rank blocks.  At depth d there are 2^(d+1) row segments of length
 * len_d = N >> (d+1); segment s covers rows [s*len_d, (s+1)*len_d).  Segment
 * s = 2n is the upper half of node n, s = 2n+1 its lower half.  Because the
 * U factor of node n's ne block lives on its upper half and the U factor of
 * its sw block on its lower half, the U factors of one level tile all N rows
 * exactly once; so do the V factors (sw.V on the upper half, ne.V on the
 * lower half).  Each level is therefore stored as two column-major N x R_d
 * arrays (column k of level d starts at uoff[d] + k*N) plus sig values
 * S[soff[d] + s*R_d + k] of the block whose U lives on segment s.  Ranks
 * below R_d are zero-padded (zero columns, zero sig).  Leaves are N/b
 * consecutive b x b column-major tiles.
 *
 * With t_s = V|_s^T x|_s, the product is
 *     (A x)|_s = D x|_s + sum_d U_d|_s (S_d[s] .* t_{s^1})
 * and with t'_s = U|_s^T x|_s the transpose is
 *     (A^T x)|_s = D^T x|_s + sum_d V_d|_s (S_d[s^1] .* t'_{s^1}).
 */
#ifndef DPLR_AMD_HIP_INTERNAL_H
#define DPLR_AMD_HIP_INTERNAL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPLRH_MAXL 24

typedef struct
{
    int N, L, b;                 /* order, depth, leaf order: N == b << L     */
    int R[DPLRH_MAXL];           /* stored (padded) rank of each level         */
    size_t uoff[DPLRH_MAXL + 1]; /* element offset of level d in U and in V    */
    size_t soff[DPLRH_MAXL + 1]; /* element offset of level d in S             */
    int *rank;                   /* true rank of the block whose U is on
                                    segment s of level d: rank[(2<<d)-2+s]     */
    double *U, *V, *S, *D;       /* host arrays in the layout above            */
} plr_flat;

/* ---- host side (C) ---- */
int plr_flat_shape(int m, const void *A, plr_flat *f, int verbose);
int plr_flat_fill(const void *A, plr_flat *f);
void plr_flat_free(plr_flat *f);

/* ---- HIP side (extern "C" from .hip files) ---- */
typedef struct dplrh_plan dplrh_plan;

/* Uploads the FP64 level arrays; returns 0 or a negative error code. */
int dplrh_plan_create(const plr_flat *f, dplrh_plan **out);
void dplrh_plan_destroy(dplrh_plan *p);
int dplrh_plan_dims(const dplrh_plan *p, int *N, int *L, int *b, int *R);
size_t dplrh_plan_bytes(const dplrh_plan *p);

/* y := alpha*op(A)*x + beta*y on device vectors (unit stride). */
int dplrh_gemv(dplrh_plan *p, int trans, double alpha, double beta,
               const double *dx, double *dy, void *stream);
/* The same with the kernels' scratch in the library's static device array
 * (no allocation; -7 when the plan needs more than DPLRX_SCRATCH_BYTES). */
int dplrh_gemv_noalloc(dplrh_plan *p, int trans, double alpha, double beta,
                       const double *dx, double *dy, void *stream);
/* C := alpha*op(A)*B + beta*C on device column-major blocks. */
/* Outer-product gradient projected onto the plan's factors (plr_grad.hip). */
int dplrh_grad_outer(dplrh_plan *p, int n, double alpha, const double *A, int lda,
                     const double *B, int ldb, double *dU, double *dV, double *dS, double *dD,
                     void *stream);
/* Batches of independent plans (plr_batch.hip): one launch per phase covers
 * every matrix, replayed from a HIP graph.  X[i], Y[i] are device
 * column-major N_i x nrhs blocks with leading dimensions ldx[i], ldy[i]. */
typedef struct dplrh_batch dplrh_batch;
int dplrh_batch_create(int count, dplrh_plan *const *plans, dplrh_batch **out);
void dplrh_batch_destroy(dplrh_batch *B);
int dplrh_batch_gemm(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                     double *const *Y, const int *ldy, void *stream);
int dplrh_batch_solve(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                      double *const *Y, const int *ldy, void *stream);
int dplrh_batch_set_tile(dplrh_batch *B, int tmax);
int dplrh_batch_tune(dplrh_batch *B, int nrhs, const double *const *X, const int *ldx,
                     double *const *Y, const int *ldy, int reps, int *tmax_out, double *ms,
                     void *stream);
/* Any precision (0 FP64, 1 FP32, 2 BF16 factors with FP32 arithmetic; FP32
 * vectors for 1 and 2) and direction; op 0 apply, op 1 solve. */
int dplrh_batch_op(dplrh_batch *B, int prec, int op, int trans, int nrhs, const void *const *X,
                   const int *ldx, void *const *Y, const int *ldy, double alpha, double beta,
                   void *stream);
/* BF16 solves refined with FP32 residuals (iters solves). */
int dplrh_batch_refined(dplrh_batch *B, int nrhs, const float *const *X, const int *ldx,
                        float *const *Y, const int *ldy, int iters, void *stream);
int dplrh_batch_tune_prec(dplrh_batch *B, int prec, int nrhs, const void *const *X,
                          const int *ldx, void *const *Y, const int *ldy, int reps,
                          int *tmax_out, double *ms, void *stream);
/* Single-precision factor copies (plr_sp.hip): type 1 FP32, 2 BF16; with
 * factor, the FP64 factorization runs first when missing. */
int dplrh_sp_factor(dplrh_plan *p, int type, int factor, int *info, void *stream);
int dplrh_sp_gemm(dplrh_plan *p, int prec, int trans, int n, float alpha, float beta,
                  const float *X, int ldx, float *Y, int ldy, void *stream);
int dplrh_sp_solve(dplrh_plan *p, int prec, int trans, int n, const float *X, int ldx, float *Y,
                   int ldy, void *stream);
int dplrh_grad_outer_f32(dplrh_plan *p, int n, float alpha, const float *A, int lda,
                         const float *B, int ldb, float *dU, float *dV, float *dS, float *dD,
                         void *stream);
int dplrh_gemm(dplrh_plan *p, int trans, int n, double alpha, double beta,
               const double *dB, int ldb, double *dC, int ldc, void *stream);

/* FP64 factored inverse (recursive Woodbury, product form); *info > 0 when
 * a leaf or capacitance matrix is singular. */
int dplrh_factor(dplrh_plan *p, int *info, void *stream);
/* X := op(A)^{-1} B on device columns, after dplrh_factor. */
int dplrh_solve(dplrh_plan *p, int trans, int n, const double *dB, int ldb, double *dX,
                int ldx, void *stream);

/* Mixed-precision factors (type 1 = f16, 2 = bf16) and refined solve:
 * `iters` low-precision solves with FP64 residuals between them; res[c]
 * (optional, host) receives max|b - A x| of column c afterwards. */
int dplrh_lp_factor(dplrh_plan *p, int type, int *info, void *stream);
int dplrh_lp_solve(dplrh_plan *p, int n, const double *dB, int ldb, double *dX, int ldx,
                   int iters, double *res, void *stream);
/* The refined solve to an error estimate <= tol (at most maxit >= 2 solves);
 * *iters and *err (optional) receive the solves made and the estimate. */
int dplrh_lp_solve_tol(dplrh_plan *p, int n, const double *dB, int ldb, double *dX, int ldx,
                       double tol, int maxit, int *iters, double *err, void *stream);
/* The same with an apply Ya = A Xa of n columns sharing the first residual
 * product's pass over the factors. */
int dplrh_lp_apply_solve_tol(dplrh_plan *p, int n, const double *dXa, int ldxa, double *dYa,
                             int ldya, const double *dB, int ldb, double *dX, int ldx, double tol,
                             int maxit, int *iters, double *err, void *stream);

/* Explicit inverse of a factored plan, downloaded per level d' into
 * hBU[d'], hBV[d'] (N*Rt[d'] doubles, [segment][column][row]) and hBsig[d']
 * (segments*Rt[d']), plus the leaves hBD.  Level d' buffers hold the own
 * columns first (R[d']) and the corrections from level d at column
 * coff[d'][d].  hnan (2^L ints, may be NULL) flags node (1 << d) - 1 + n when
 * its Woodbury step normalised a zero column with a zero threshold. */
int dplrh_geinv(dplrh_plan *p, double eps, const int *Rt, int coff[][DPLRH_MAXL],
                double **hBU, double **hBV, double **hBsig, double *hBD, int *hnan,
                void *stream);
/* Rank-k update data: per level d the k appended columns of every block
 * ([segment][column][row], scaled as dplrgeprk stores them) and their sig;
 * the leaves hD are updated in place. */
int dplrh_geprk(int N, int L, int b, int k, double alpha, double eps, const double *hu,
                const double *hv, const double *hs, double *hD, double **hUn, double **hVn,
                double **hsig, void *stream);
/* on != 0: dplrh_geprk calls of this thread keep their device buffers for
 * the next call (grown as needed); 0: release them and allocate per call. */
void dplrh_prk_pool(int on);
/* The same for the host buffers of dplrgeprk, and dplrh_prk_pool with it
 * (host layer, dplr_inv.c): dplrgeppmm turns it on around its corrections. */
void plr_prk_pool(int on);

/* Level products of dplrgeppmm: for every level d with R_d > 0,
 * PU[d] := alpha * op_ta(A_{>d}) * (tb ? V_d(B) : U_d(B))   (N x R_d(B)) and
 * PV[d] := alpha * op_tb(B_{>d})^T * (ta ? U_d(A) : V_d(A)) (N x R_d(A)),
 * host column-major arrays (ld N); A_{>d} is the block-diagonal forest of
 * the subtrees below depth d (levels d+1 .. L-1 and the leaves). */
int dplrh_ppmm_levels(dplrh_plan *pa, dplrh_plan *pb, int ta, int tb, double alpha, double **PU,
                      double **PV, void *stream);

/* Device memory helpers for the drop-in entry points. */
void *dplrh_dev_alloc(size_t bytes);
void dplrh_dev_free(void *p);
int dplrh_h2d(void *dst, const void *src, size_t bytes, void *stream);
int dplrh_d2h(void *dst, const void *src, size_t bytes, void *stream);
int dplrh_sync(void *stream);
/* Strided gather/scatter and transpose on device. */
int dplrh_gather(const double *src, int inc, int n, double *dst, void *stream);
int dplrh_transpose(int rows, int cols, const double *src, int lds, double *dst,
                    int ldd, void *stream);
/* y := beta*beta*y (or 0 when beta == 0): the alpha == 0 quirk of gepmv. */
int dplrh_scale2(int n, int cols, double beta, double *dy, int ldy, void *stream);
/* C := beta*C (or 0 when beta == 0). */
int dplrh_scale(int n, int cols, double beta, double *dy, int ldy, void *stream);
const char *dplrh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
