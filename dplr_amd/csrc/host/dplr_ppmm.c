/*
 * dplrgeppmm: C := alpha*op(A)*op(B) + beta*C for three PLR matrices of one
 * uniform structure (reference src/dplrppmm.c:998 and its workers
 * dplrppmm/dplrpptmm/dplrptpmm/dplrptptmm, :492 .. :868).
 *
 * The output is a PLR matrix whose ranks grow: every off-diagonal block of C
 * takes [beta * its own columns | the products with the subtrees | the
 * low-rank x low-rank corrections of its ancestors], normalised with the
 * reference's per-level threshold (eps/4 per level), so the order of the
 * operations is observable and the recursion below keeps the reference's.
 * The arithmetic that dominates -- every product of a diagonal subtree with
 * an off-diagonal block's factors, which the reference forms call by call
 * with dplrgepdmm -- is done on the GPU, one product per level of the tree
 * (dplrh_ppmm_levels); the recursion takes its columns from those.  The
 * corrections are added into the subtrees by dplrgeprk (GPU).
 *
 * A transposed operand is handled as a virtual PLR matrix: the ne block of
 * X^T is (X.sw)^T, i.e. u and v of X.sw swapped with the same sig, its sw
 * block is (X.ne)^T, its diagonal blocks are the transposed subtrees, and
 * its leaves the transposed leaves.  The four workers of the reference are
 * the non-transposed one applied to op(A), op(B).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dplr.h"
#include "../dplr_hip.h"

/* an operand op(X): X itself (t = 0) or its transpose (t = 1) */
typedef struct
{
    const DPlrMtx *X;
    int t;
} opx;

/* a low-rank block of op(X): u * diag(sig) * v^T, columns of length size */
typedef struct
{
    int rank;
    const double *u, *v, *sig;
} lrv;

static lrv lr_as(const DPlrMtx *B, int t)
{
    lrv r = {B->dlr.rank, t ? B->dlr.v : B->dlr.u, t ? B->dlr.u : B->dlr.v, B->dlr.sig};
    return r;
}
static lrv op_ne(opx a) { return a.t ? lr_as(a.X->dh.sw, 1) : lr_as(a.X->dh.ne, 0); }
static lrv op_sw(opx a) { return a.t ? lr_as(a.X->dh.ne, 1) : lr_as(a.X->dh.sw, 0); }
static opx op_nw(opx a) { opx r = {a.X->dh.nw, a.t}; return r; }
static opx op_se(opx a) { opx r = {a.X->dh.se, a.t}; return r; }

typedef struct
{
    int N;          /* order of the whole matrix (leading dimension of PU, PV) */
    double alpha;
    double **PU;    /* per level: alpha * op(A)_{>d} * U_d(op(B)), N x R_d */
    double **PV;    /* per level: alpha * op(B)_{>d}^T * V_d(op(A)), N x R_d */
} ppctx;

/* C (a low-rank block) := beta * C: sig scaled by |beta|, the sign of a
 * negative beta put into u (src/dplrppmm.c:216-238). */
static void lr_scale(DPlrMtx *C, int m, double beta)
{
    if (beta == 0.0 || C->dlr.rank == 0) {
        C->dlr.rank = 0;
        return;
    }
    const int r = C->dlr.rank;
    if (beta > 0.0) {
        for (int k = 0; k < r; ++k) C->dlr.sig[k] *= beta;
    } else {
        for (int k = 0; k < r; ++k) C->dlr.sig[k] *= -beta;
        const size_t n = (size_t)m * r;
        for (size_t i = 0; i < n; ++i) C->dlr.u[i] = -C->dlr.u[i];
    }
}

/* Columns just written (k of length m): each is normalised against the
 * threshold, sig = |col| * s[j], or sig = 0 with the column left as it is
 * when that is below eps (src/dplrppmm.c:246-258). */
/* Work (m * columns) below which the host loops here stay on one thread. */
#define PPMM_PAR ((size_t)1 << 15)

static void lr_normalise(double *cols, int m, int k, const double *s, double *sig, double eps)
{
#pragma omp parallel for if ((size_t)m * k >= PPMM_PAR)
    for (int j = 0; j < k; ++j) {
        double *c = cols + (size_t)j * m;
        double ss = 0.0;
        for (int i = 0; i < m; ++i) ss += c[i] * c[i];
        const double nrm = sqrt(ss);
        if (nrm * s[j] < eps) {
            sig[j] = 0.0;
        } else {
            sig[j] = nrm * s[j];
            const double f = 1.0 / nrm;
            for (int i = 0; i < m; ++i) c[i] *= f;
        }
    }
}

static int lr_check(const char *fn, const DPlrMtx *C, int m)
{
    if (C->type != PLR_LRMTX || C->size != m) {
        fprintf(stderr, "%s: invalid C block\n", fn);
        return -1;
    }
    return 0;
}

/* C := alpha * op(A)_sub * B' + beta * C, B' a low-rank block of op(B):
 * new u = the product (rows of PU, leading dimension ld), v = B'.v
 * (src/dplrppmm.c:183, dplrplmm). */
static void plmm(DPlrMtx *C, int m, lrv b, const double *pu, int ld, double alpha, double beta,
                 double eps)
{
    if (lr_check("dplrgeppmm", C, m)) return;
    lr_scale(C, m, beta);
    if (alpha == 0.0 || b.rank == 0) return;
    if (C->dlr.capacity < C->dlr.rank + b.rank) dplrresizelr(C, C->dlr.rank + b.rank);
    const int r = C->dlr.rank;
    for (int j = 0; j < b.rank; ++j)
        memcpy(C->dlr.u + (size_t)(r + j) * m, pu + (size_t)j * ld, (size_t)m * sizeof(double));
    memcpy(C->dlr.v + (size_t)r * m, b.v, (size_t)m * b.rank * sizeof(double));
    lr_normalise(C->dlr.u + (size_t)r * m, m, b.rank, b.sig, C->dlr.sig + r, eps);
    C->dlr.rank += b.rank;
}

/* C := alpha * A' * op(B)_sub + beta * C, A' a low-rank block of op(A):
 * new v = op(B)_sub^T A'.v (rows of PV), u = A'.u (src/dplrppmm.c:298,
 * dplrlpmm). */
static void lpmm(DPlrMtx *C, int m, lrv a, const double *pv, int ld, double alpha, double beta,
                 double eps)
{
    if (lr_check("dplrgeppmm", C, m)) return;
    lr_scale(C, m, beta);
    if (alpha == 0.0 || a.rank == 0) return;
    if (C->dlr.capacity < C->dlr.rank + a.rank) dplrresizelr(C, C->dlr.rank + a.rank);
    const int r = C->dlr.rank;
    for (int j = 0; j < a.rank; ++j)
        memcpy(C->dlr.v + (size_t)(r + j) * m, pv + (size_t)j * ld, (size_t)m * sizeof(double));
    memcpy(C->dlr.u + (size_t)r * m, a.u, (size_t)m * a.rank * sizeof(double));
    lr_normalise(C->dlr.v + (size_t)r * m, m, a.rank, a.sig, C->dlr.sig + r, eps);
    C->dlr.rank += a.rank;
}

/* The product of two low-rank blocks as one of rank min(r1, r2), into
 * u3, s3, v3 (m x rank); returns the rank (src/dplrppmm.c:25, dplrgellmm,
 * with beta = 0 and no previous columns). */
static int llmm(int m, lrv a, lrv b, double alpha, double *u3, double *s3, double *v3, double eps)
{
    const int r1 = a.rank, r2 = b.rank;
    if (m == 0 || r1 == 0 || r2 == 0 || alpha == 0.0) return 0;
    if (r1 < r2) {
        /* temp (r2 x r1) = s2 .* (u2^T v1); u3 = u1, v3 = alpha v2 temp */
        double *temp = malloc((size_t)r1 * r2 * sizeof(double));
        if (!temp) return 0;
#pragma omp parallel for collapse(2) if ((size_t)m * r1 * r2 >= PPMM_PAR)
        for (int j = 0; j < r1; ++j)
            for (int i = 0; i < r2; ++i) {
                const double *x = b.u + (size_t)i * m, *y = a.v + (size_t)j * m;
                double s = 0.0;
                for (int l = 0; l < m; ++l) s += x[l] * y[l];
                temp[i + (size_t)j * r2] = b.sig[i] * s;
            }
        memcpy(u3, a.u, (size_t)m * r1 * sizeof(double));
#pragma omp parallel for if ((size_t)m * r1 * r2 >= PPMM_PAR)
        for (int j = 0; j < r1; ++j) {
            double *c = v3 + (size_t)j * m;
            for (int l = 0; l < m; ++l) c[l] = 0.0;
            for (int i = 0; i < r2; ++i) {
                const double t = alpha * temp[i + (size_t)j * r2];
                const double *x = b.v + (size_t)i * m;
                for (int l = 0; l < m; ++l) c[l] += x[l] * t;
            }
        }
        free(temp);
        lr_normalise(v3, m, r1, a.sig, s3, eps);
        return r1;
    }
    /* temp (r1 x r2) = s1 .* (v1^T u2); v3 = v2, u3 = alpha u1 temp */
    double *temp = malloc((size_t)r1 * r2 * sizeof(double));
    if (!temp) return 0;
#pragma omp parallel for collapse(2) if ((size_t)m * r1 * r2 >= PPMM_PAR)
    for (int j = 0; j < r2; ++j)
        for (int i = 0; i < r1; ++i) {
            const double *x = a.v + (size_t)i * m, *y = b.u + (size_t)j * m;
            double s = 0.0;
            for (int l = 0; l < m; ++l) s += x[l] * y[l];
            temp[i + (size_t)j * r1] = a.sig[i] * s;
        }
    memcpy(v3, b.v, (size_t)m * r2 * sizeof(double));
#pragma omp parallel for if ((size_t)m * r1 * r2 >= PPMM_PAR)
    for (int j = 0; j < r2; ++j) {
        double *c = u3 + (size_t)j * m;
        for (int l = 0; l < m; ++l) c[l] = 0.0;
        for (int i = 0; i < r1; ++i) {
            const double t = alpha * temp[i + (size_t)j * r1];
            const double *x = a.u + (size_t)i * m;
            for (int l = 0; l < m; ++l) c[l] += x[l] * t;
        }
    }
    free(temp);
    lr_normalise(u3, m, r2, b.sig, s3, eps);
    return r2;
}

/* dense leaves: C := alpha op(A) op(B) + beta C (C not read when beta = 0) */
static void dense_mm(int m, double alpha, opx a, opx b, double beta, double *C)
{
    const double *A = a.X->dd.d, *B = b.X->dd.d;
    double *col = malloc((size_t)m * sizeof(double));
    if (!col) return;
    for (int j = 0; j < m; ++j) {
        for (int i = 0; i < m; ++i) col[i] = 0.0;
        if (alpha != 0.0)
            for (int k = 0; k < m; ++k) {
                const double bkj = b.t ? B[j + (size_t)k * m] : B[k + (size_t)j * m];
                if (a.t) {
                    for (int i = 0; i < m; ++i) col[i] += A[k + (size_t)i * m] * bkj;
                } else {
                    const double *ak = A + (size_t)k * m;
                    for (int i = 0; i < m; ++i) col[i] += ak[i] * bkj;
                }
            }
        double *c = C + (size_t)j * m;
        for (int i = 0; i < m; ++i) c[i] = alpha * col[i] + (beta == 0.0 ? 0.0 : beta * c[i]);
    }
    free(col);
}

static void ppmm_rec(const ppctx *cx, int d, int row0, int m, opx a, opx b, double beta,
                     DPlrMtx *C, double eps)
{
    const DPlrMtx *A = a.X, *B = b.X;
    if (A->type != B->type || B->type != C->type) {
        fprintf(stderr, "dplrgeppmm: A, B and C type do not match\n");
        return;
    }
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "dplrgeppmm: invalid A, B or C type\n");
        return;
    }
    if (A->size != m || B->size != m || C->size != m) {
        fprintf(stderr, "dplrgeppmm: invalid A, B or C size\n");
        return;
    }
    if (m == 0) return;
    const double alpha = cx->alpha;
    if (A->type == PLR_DMTX) {
        dense_mm(m, alpha, a, b, beta, C->dd.d);
        return;
    }
    const double eps2 = eps / 4.0, one = 1.0;
    const int h = m / 2, ld = cx->N;
    const double *PU = cx->PU[d], *PV = cx->PV[d];
    const lrv ane = op_ne(a), asw = op_sw(a), bne = op_ne(b), bsw = op_sw(b);
    /* C.ne := alpha (A'.nw B'.ne + A'.ne B'.se) + beta C.ne */
    plmm(C->dh.ne, h, bne, PU ? PU + row0 : NULL, ld, alpha, beta, eps2);
    lpmm(C->dh.ne, h, ane, PV ? PV + row0 + h : NULL, ld, alpha, 1.0, eps2);
    /* C.sw := alpha (A'.sw B'.nw + A'.se B'.sw) + beta C.sw */
    lpmm(C->dh.sw, h, asw, PV ? PV + row0 : NULL, ld, alpha, beta, eps2);
    plmm(C->dh.sw, h, bsw, PU ? PU + row0 + h : NULL, ld, alpha, 1.0, eps2);
    /* corrections of the diagonal blocks: rank min(r1, r2) */
    const int r1 = ane.rank < bsw.rank ? ane.rank : bsw.rank;
    const int r2 = asw.rank < bne.rank ? asw.rank : bne.rank;
    const int r = r1 > r2 ? r1 : r2;
    double *u = malloc((size_t)h * (r ? r : 1) * sizeof(double));
    double *v = malloc((size_t)h * (r ? r : 1) * sizeof(double));
    double *s = malloc((size_t)(r ? r : 1) * sizeof(double));
    if (!u || !v || !s) {
        fprintf(stderr, "dplrgeppmm: memory allocation failed\n");
        free(u);
        free(v);
        free(s);
        return;
    }
    /* C.nw := alpha A'.nw B'.nw + beta C.nw, then + alpha A'.ne B'.sw */
    ppmm_rec(cx, d + 1, row0, h, op_nw(a), op_nw(b), beta, C->dh.nw, eps2);
    int k = llmm(h, ane, bsw, alpha, u, s, v, eps2);
    dplrgeprk(h, k, (double *)&one, u, h, v, h, s, 1, C->dh.nw, eps2);
    /* C.se := alpha A'.se B'.se + beta C.se, then + alpha A'.sw B'.ne */
    ppmm_rec(cx, d + 1, row0 + h, h, op_se(a), op_se(b), beta, C->dh.se, eps2);
    k = llmm(h, asw, bne, alpha, u, s, v, eps2);
    dplrgeprk(h, k, (double *)&one, u, h, v, h, s, 1, C->dh.se, eps2);
    free(u);
    free(v);
    free(s);
}

static int make_plan_of(int m, const DPlrMtx *A, dplrh_plan **p)
{
    plr_flat f;
    int rc = plr_flat_shape(m, A, &f, 1);
    if (rc) return rc;
    rc = plr_flat_fill(A, &f);
    if (!rc) rc = dplrh_plan_create(&f, p);
    plr_flat_free(&f);
    return rc;
}

/* reference src/dplrppmm.c:998 */
void dplrgeppmm(char transa, char transb, int m, double *alpha, DPlrMtx *A, DPlrMtx *B,
                double *beta, DPlrMtx *C, double eps)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size %d\n", __func__, m);
        return;
    }
    if (A->type != B->type || B->type != C->type) {
        fprintf(stderr, "%s: A, B and C type do not match\n", __func__);
        return;
    }
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "%s: invalid A, B or C type\n", __func__);
        return;
    }
    if (A->size != m || B->size != m || C->size != m) {
        fprintf(stderr, "%s: invalid A, B or C size\n", __func__);
        return;
    }
    if (m == 0) return;
    if (C == A || C == B) {
        fprintf(stderr, "%s: C must not be A or B\n", __func__);
        return;
    }
    const int ta = !(transa == 'N' || transa == 'n'), tb = !(transb == 'N' || transb == 'n');
    opx a = {A, ta}, b = {B, tb};
    ppctx cx;
    memset(&cx, 0, sizeof(cx));
    cx.N = m;
    cx.alpha = *alpha;
    double *PU[DPLRH_MAXL] = {0}, *PV[DPLRH_MAXL] = {0};
    cx.PU = PU;
    cx.PV = PV;
    dplrh_plan *pa = NULL, *pb = NULL;
    int rc = 0;
    if (A->type == PLR_HMTX && *alpha != 0.0) {
        if ((rc = make_plan_of(m, A, &pa)) || (rc = make_plan_of(m, B, &pb))) {
            fprintf(stderr, "%s: device upload failed: %s\n", __func__, dplrh_last_error());
            dplrh_plan_destroy(pa);
            dplrh_plan_destroy(pb);
            return;
        }
        int N, L, lb, Ra[DPLRH_MAXL], Rb[DPLRH_MAXL];
        dplrh_plan_dims(pa, &N, &L, &lb, Ra);
        dplrh_plan_dims(pb, &N, &L, &lb, Rb);
        for (int d = 0; d < L && !rc; ++d) {
            PU[d] = malloc((size_t)m * (Rb[d] ? Rb[d] : 1) * sizeof(double));
            PV[d] = malloc((size_t)m * (Ra[d] ? Ra[d] : 1) * sizeof(double));
            if (!PU[d] || !PV[d]) rc = -3;
        }
        if (!rc) rc = dplrh_ppmm_levels(pa, pb, ta, tb, *alpha, PU, PV, NULL);
        dplrh_plan_destroy(pa);
        dplrh_plan_destroy(pb);
        if (rc) fprintf(stderr, "%s: GPU execution failed (%d): %s\n", __func__, rc,
                        dplrh_last_error());
    }
    if (!rc) {
        /* the corrections reuse their buffers from one node to the next */
        plr_prk_pool(1);
        ppmm_rec(&cx, 0, 0, m, a, b, *beta, C, eps);
        plr_prk_pool(0);
    }
    for (int d = 0; d < DPLRH_MAXL; ++d) {
        free(PU[d]);
        free(PV[d]);
    }
}
