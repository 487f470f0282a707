/*
 * Drop-in entry points of dplr.h that run on the GPU, and the dplrx
 * device-resident API.  Argument checks and their messages follow the
 * reference workers; the arithmetic is done by the HIP layer.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dplr.h"
#include "dplrx.h"
#include "../dplr_hip.h"

static int is_n(char c) { return c == 'N' || c == 'n'; }

/* Checks of the reference workers (e.g. src/dplrpmv.c:116-130). */
static int check_root(const char *fn, int m, const DPlrMtx *A)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size %d\n", fn, m);
        return -1;
    }
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "%s: invalid A type\n", fn);
        return -1;
    }
    if (A->size != m) {
        fprintf(stderr, "%s: invalid A size\n", fn);
        return -1;
    }
    return 0;
}

static int make_plan(int m, const DPlrMtx *A, dplrh_plan **plan)
{
    plr_flat f;
    int rc = plr_flat_shape(m, A, &f, 1);
    if (rc) return rc;
    rc = plr_flat_fill(A, &f);
    if (rc) return rc;
    rc = dplrh_plan_create(&f, plan);
    plr_flat_free(&f);
    if (rc) fprintf(stderr, "dplr: device upload failed: %s\n", dplrh_last_error());
    return rc;
}

static void report(const char *fn, int rc)
{
    if (rc) fprintf(stderr, "%s: GPU execution failed (%d): %s\n", fn, rc, dplrh_last_error());
}

/* reference src/dplrpmv.c:262 */
void dplrgepmv(char trans, int m, double *alpha, DPlrMtx *A, double *x, int incx,
               double *beta, double *y, int incy)
{
    if (check_root(__func__, m, A)) return;
    if (incx <= 0 || incy <= 0) {
        fprintf(stderr, "%s: incx and incy must be positive\n", __func__);
        return;
    }
    if (m == 0) return;
    dplrh_plan *p = NULL;
    if (make_plan(m, A, &p)) return;
    const size_t bytes = (size_t)m * sizeof(double);
    double *hx = malloc(bytes), *hy = malloc(bytes);
    double *dx = dplrh_dev_alloc(bytes), *dy = dplrh_dev_alloc(bytes);
    int rc = -3;
    if (hx && hy && dx && dy) {
        for (int i = 0; i < m; ++i) hx[i] = x[(size_t)i * incx];
        /* y is read when beta != 0, and by the beta^2 quirk when alpha == 0 */
        for (int i = 0; i < m; ++i) hy[i] = y[(size_t)i * incy];
        rc = dplrh_h2d(dx, hx, bytes, NULL);
        if (!rc) rc = dplrh_h2d(dy, hy, bytes, NULL);
        if (!rc) rc = dplrh_gemv(p, !is_n(trans), *alpha, *beta, dx, dy, NULL);
        if (!rc) rc = dplrh_d2h(hy, dy, bytes, NULL);
        if (!rc)
            for (int i = 0; i < m; ++i) y[(size_t)i * incy] = hy[i];
    }
    report(__func__, rc);
    dplrh_dev_free(dx);
    dplrh_dev_free(dy);
    free(hx);
    free(hy);
    dplrh_plan_destroy(p);
}

/* Upload a rows x cols column-major host block with leading dimension ld. */
static int upload_block(double *dst, const double *src, int rows, int cols, int ld)
{
    if (ld == rows) return dplrh_h2d(dst, src, (size_t)rows * cols * sizeof(double), NULL);
    for (int j = 0; j < cols; ++j) {
        int rc = dplrh_h2d(dst + (size_t)j * rows, src + (size_t)j * ld,
                           (size_t)rows * sizeof(double), NULL);
        if (rc) return rc;
    }
    return 0;
}

static int download_block(double *dst, const double *src, int rows, int cols, int ld)
{
    if (ld == rows) return dplrh_d2h(dst, src, (size_t)rows * cols * sizeof(double), NULL);
    double *h = malloc((size_t)rows * cols * sizeof(double));
    if (!h) return -3;
    int rc = dplrh_d2h(h, src, (size_t)rows * cols * sizeof(double), NULL);
    if (!rc)
        for (int j = 0; j < cols; ++j)
            memcpy(dst + (size_t)j * ld, h + (size_t)j * rows, (size_t)rows * sizeof(double));
    free(h);
    return rc;
}

/* reference src/dplrpdmm.c:454 */
void dplrgepdmm(char transa, char transb, int m, int n, double *alpha, DPlrMtx *A,
                double *B, int ldb, double *beta, double *C, int ldc)
{
    if (check_root(__func__, m, A)) return;
    if (ldc < m) {
        fprintf(stderr, "%s: invalid leading dimensions\n", __func__);
        return;
    }
    if (n < 0) {
        fprintf(stderr, "%s: invalid number of columns %d\n", __func__, n);
        return;
    }
    if (m == 0 || n == 0) return;
    const int tb = !is_n(transb);
    if (*alpha != 0.0 && ldb < (tb ? n : m)) {
        fprintf(stderr, "%s: invalid leading dimension of B\n", __func__);
        return;
    }
    dplrh_plan *p = NULL;
    if (make_plan(m, A, &p)) return;
    const size_t mn = (size_t)m * n;
    double *dB = dplrh_dev_alloc(mn * sizeof(double));
    double *dT = tb ? dplrh_dev_alloc(mn * sizeof(double)) : NULL;
    double *dC = dplrh_dev_alloc(mn * sizeof(double));
    int rc = -3;
    if (dB && dC && (!tb || dT)) {
        rc = 0;
        if (*alpha != 0.0) {
            if (!tb) {
                rc = upload_block(dB, B, m, n, ldb);
            } else {
                rc = upload_block(dT, B, n, m, ldb);
                if (!rc) rc = dplrh_transpose(m, n, dT, n, dB, m, NULL);
            }
        }
        if (!rc && *beta != 0.0) rc = upload_block(dC, C, m, n, ldc);
        if (!rc) rc = dplrh_gemm(p, !is_n(transa), n, *alpha, *beta, dB, m, dC, m, NULL);
        if (!rc) rc = download_block(C, dC, m, n, ldc);
    }
    report(__func__, rc);
    dplrh_dev_free(dB);
    dplrh_dev_free(dT);
    dplrh_dev_free(dC);
    dplrh_plan_destroy(p);
}

/* ---- dplrx ---------------------------------------------------------------- */

int dplrx_plan_create(int m, DPlrMtx *A, dplrx_plan **plan)
{
    *plan = NULL;
    if (check_root(__func__, m, A)) return -1;
    return make_plan(m, A, plan);
}

void dplrx_plan_destroy(dplrx_plan *plan) { dplrh_plan_destroy(plan); }

int dplrx_plan_info(const dplrx_plan *plan, int *m, int *nlvl, int *leaf, int *ranks)
{
    int R[DPLRH_MAXL];
    dplrh_plan_dims(plan, m, nlvl, leaf, R);
    if (ranks)
        for (int d = 0; d < *nlvl; ++d) ranks[d] = R[d];
    return 0;
}

size_t dplrx_plan_bytes(const dplrx_plan *plan) { return dplrh_plan_bytes(plan); }

int dplrx_gepmv(dplrx_plan *plan, char trans, double alpha, const double *dx, double beta,
                double *dy, void *stream)
{
    return dplrh_gemv(plan, !is_n(trans), alpha, beta, dx, dy, stream);
}

int dplrx_gepmv_noalloc(dplrx_plan *plan, char trans, double alpha, const double *dx,
                        double beta, double *dy, void *stream)
{
    return dplrh_gemv_noalloc(plan, !is_n(trans), alpha, beta, dx, dy, stream);
}

int dplrx_gepdmm(dplrx_plan *plan, char transa, int n, double alpha, const double *dB,
                 int ldb, double beta, double *dC, int ldc, void *stream)
{
    int N, L, b, R[DPLRH_MAXL];
    dplrh_plan_dims(plan, &N, &L, &b, R);
    if (n < 0 || ldb < N || ldc < N) return -1;
    return dplrh_gemm(plan, !is_n(transa), n, alpha, beta, dB, ldb, dC, ldc, stream);
}

int dplrx_factor(dplrx_plan *plan, int *info, void *stream)
{
    return dplrh_factor(plan, info, stream);
}

int dplrx_solve(dplrx_plan *plan, char trans, int n, const double *dB, int ldb, double *dX,
                int ldx, void *stream)
{
    int N, L, b, R[DPLRH_MAXL];
    dplrh_plan_dims(plan, &N, &L, &b, R);
    if (n < 0 || ldb < N || ldx < N || (n > 0 && dB == dX)) return -1;
    return dplrh_solve(plan, !is_n(trans), n, dB, ldb, dX, ldx, stream);
}

int dplrx_factor_mixed(dplrx_plan *plan, int type, int *info, void *stream)
{
    return dplrh_lp_factor(plan, type, info, stream);
}

int dplrx_solve_refined(dplrx_plan *plan, int n, const double *dB, int ldb, double *dX,
                        int ldx, int iters, double *res, void *stream)
{
    int N, L, b, R[DPLRH_MAXL];
    dplrh_plan_dims(plan, &N, &L, &b, R);
    if (n < 0 || iters < 1 || ldb < N || ldx < N || (n > 0 && dB == dX)) return -1;
    return dplrh_lp_solve(plan, n, dB, ldb, dX, ldx, iters, res, stream);
}

int dplrx_solve_refined_tol(dplrx_plan *plan, int n, const double *dB, int ldb, double *dX,
                            int ldx, double tol, int maxit, int *iters, double *err, void *stream)
{
    int N, L, b, R[DPLRH_MAXL];
    dplrh_plan_dims(plan, &N, &L, &b, R);
    if (n < 0 || maxit < 2 || !(tol >= 0.0) || ldb < N || ldx < N || (n > 0 && dB == dX))
        return -1;
    return dplrh_lp_solve_tol(plan, n, dB, ldb, dX, ldx, tol, maxit, iters, err, stream);
}

int dplrx_apply_solve_refined_tol(dplrx_plan *plan, int n, const double *dXa, int ldxa,
                                  double *dYa, int ldya, const double *dB, int ldb, double *dX,
                                  int ldx, double tol, int maxit, int *iters, double *err,
                                  void *stream)
{
    int N, L, b, R[DPLRH_MAXL];
    dplrh_plan_dims(plan, &N, &L, &b, R);
    if (n < 0 || maxit < 2 || !(tol >= 0.0) || ldb < N || ldx < N || ldxa < N || ldya < N ||
        (n > 0 && (dB == dX || dXa == dYa || dYa == dX || dYa == dB)))
        return -1;
    return dplrh_lp_apply_solve_tol(plan, n, dXa, ldxa, dYa, ldya, dB, ldb, dX, ldx, tol, maxit,
                                    iters, err, stream);
}

int dplrx_factor_single(dplrx_plan *plan, int type, int *info, void *stream)
{
    return dplrh_sp_factor(plan, type, 1, info, stream);
}

int dplrx_gepdmm_single(dplrx_plan *plan, int type, char trans, int n, float alpha,
                        const float *dB, int ldb, float beta, float *dC, int ldc, void *stream)
{
    if (n < 0 || (type != 1 && type != 2)) return -1;
    return dplrh_sp_gemm(plan, type, !is_n(trans), n, alpha, beta, dB, ldb, dC, ldc, stream);
}

int dplrx_solve_single(dplrx_plan *plan, int type, char trans, int n, const float *dB, int ldb,
                       float *dX, int ldx, void *stream)
{
    if (n < 0 || (type != 1 && type != 2) || (n > 0 && dB == dX)) return -1;
    return dplrh_sp_solve(plan, type, !is_n(trans), n, dB, ldb, dX, ldx, stream);
}

const char *dplrx_last_error(void) { return dplrh_last_error(); }

/* ---- batches ------------------------------------------------------------ */

int dplrx_batch_create(int count, dplrx_plan *const *plans, dplrx_batch **batch)
{
    return dplrh_batch_create(count, plans, batch);
}

void dplrx_batch_destroy(dplrx_batch *batch) { dplrh_batch_destroy(batch); }

int dplrx_batch_gepdmm(dplrx_batch *batch, int nrhs, const double *const *dX, const int *ldx,
                       double *const *dY, const int *ldy, void *stream)
{
    return dplrh_batch_gemm(batch, nrhs, dX, ldx, dY, ldy, stream);
}

int dplrx_batch_solve(dplrx_batch *batch, int nrhs, const double *const *dB, const int *ldb,
                      double *const *dX, const int *ldx, void *stream)
{
    return dplrh_batch_solve(batch, nrhs, dB, ldb, dX, ldx, stream);
}

int dplrx_batch_tune(dplrx_batch *batch, int nrhs, const double *const *dX, const int *ldx,
                     double *const *dY, const int *ldy, int reps, int *tile, double *ms,
                     void *stream)
{
    return dplrh_batch_tune(batch, nrhs, dX, ldx, dY, ldy, reps, tile, ms, stream);
}

int dplrx_batch_set_tile(dplrx_batch *batch, int tile) { return dplrh_batch_set_tile(batch, tile); }

int dplrx_batch_op(dplrx_batch *batch, int prec, int op, char trans, int nrhs,
                   const void *const *dX, const int *ldx, void *const *dY, const int *ldy,
                   double alpha, double beta, void *stream)
{
    return dplrh_batch_op(batch, prec, op, !is_n(trans), nrhs, dX, ldx, dY, ldy, alpha, beta,
                          stream);
}

int dplrx_batch_solve_refined(dplrx_batch *batch, int nrhs, const float *const *dB,
                              const int *ldb, float *const *dX, const int *ldx, int iters,
                              void *stream)
{
    return dplrh_batch_refined(batch, nrhs, dB, ldb, dX, ldx, iters, stream);
}

int dplrx_batch_tune_prec(dplrx_batch *batch, int prec, int nrhs, const void *const *dX,
                          const int *ldx, void *const *dY, const int *ldy, int reps, int *tile,
                          double *ms, void *stream)
{
    return dplrh_batch_tune_prec(batch, prec, nrhs, dX, ldx, dY, ldy, reps, tile, ms, stream);
}

/* ---- adjoints ----------------------------------------------------------- */

int dplrx_grad_outer(dplrx_plan *plan, int n, double alpha, const double *dGa, int lda,
                     const double *dGb, int ldb, double *dU, double *dV, double *dS, double *dD,
                     void *stream)
{
    return dplrh_grad_outer(plan, n, alpha, dGa, lda, dGb, ldb, dU, dV, dS, dD, stream);
}

int dplrx_grad_outer_f32(dplrx_plan *plan, int n, float alpha, const float *dGa, int lda,
                         const float *dGb, int ldb, float *dU, float *dV, float *dS, float *dD,
                         void *stream)
{
    return dplrh_grad_outer_f32(plan, n, alpha, dGa, lda, dGb, ldb, dU, dV, dS, dD, stream);
}
