/*
 * Node lifecycle and small host helpers of the drop-in API.
 *
 * Allocation semantics follow the reference: buffers come from malloc and
 * are NaN-filled so that a read of an unwritten slot shows up as NaN
 * (reference src/dplrini.c:58-62, :109-121); capacity growth at least
 * doubles (reference src/dplrresizelr.c:54).  Sizes are computed in size_t,
 * so orders above 46340 do not overflow as the reference's int products do.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dplr.h"

static void fill_nan(double *p, size_t n)
{
    const double q = nan("");
#pragma omp parallel for schedule(static) if (n >= ((size_t)1 << 18))
    for (size_t i = 0; i < n; ++i) p[i] = q;
}

/* reference src/dplrini.c:18 */
void dplrini(DPlrMtx *A)
{
    memset(A, 0, sizeof(*A));
    A->type = PLR_UMTX;
}

/* reference src/dplrini.c:35 */
void dplrinid(int m, DPlrMtx *A)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size\n", __func__);
        return;
    }
    A->type = PLR_DMTX;
    A->size = m;
    A->dd.d = NULL;
    if (m == 0) return;
    A->dd.d = malloc((size_t)m * (size_t)m * sizeof(double));
    if (!A->dd.d) {
        fprintf(stderr, "%s: memory allocation failed\n", __func__);
        return;
    }
    fill_nan(A->dd.d, (size_t)m * (size_t)m);
}

/* reference src/dplrini.c:80 */
void dplrinilr(int m, DPlrMtx *A, int capacity)
{
    if (m < 0 || capacity < 0) {
        fprintf(stderr, "%s: invalid size or capacity\n", __func__);
        return;
    }
    A->type = PLR_LRMTX;
    A->size = m;
    A->dlr.rank = 0;
    A->dlr.capacity = capacity;
    A->dlr.u = A->dlr.v = A->dlr.sig = NULL;
    if (capacity == 0) return;
    size_t n = (size_t)m * (size_t)capacity;
    A->dlr.u = malloc(n * sizeof(double));
    A->dlr.v = malloc(n * sizeof(double));
    A->dlr.sig = malloc((size_t)capacity * sizeof(double));
    if (!A->dlr.u || !A->dlr.v || !A->dlr.sig) {
        fprintf(stderr, "%s: memory allocation failed\n", __func__);
        return;
    }
    fill_nan(A->dlr.u, n);
    fill_nan(A->dlr.v, n);
    fill_nan(A->dlr.sig, (size_t)capacity);
}

/* reference src/dplrini.c:142 */
void dplrinih(int m, DPlrMtx *A)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid size\n", __func__);
        return;
    }
    if (m % 2 != 0) {
        fprintf(stderr, "%s: m must be even\n", __func__);
        return;
    }
    DPlrMtx *c[4];
    for (int i = 0; i < 4; ++i) {
        c[i] = malloc(sizeof(DPlrMtx));
        if (!c[i]) {
            fprintf(stderr, "%s: memory allocation failed\n", __func__);
            for (int j = 0; j < i; ++j) free(c[j]);
            return;
        }
        dplrini(c[i]);
    }
    A->type = PLR_HMTX;
    A->size = m;
    A->dh.nw = c[0];
    A->dh.ne = c[1];
    A->dh.se = c[2];
    A->dh.sw = c[3];
}

/* reference src/dplrini.c:183.  Children allocated by dplrinih are freed as
 * well and the node is reset to UMTX, so clearing twice is harmless. */
void dplrclr(DPlrMtx *A)
{
    switch (A->type) {
    case PLR_DMTX:
        free(A->dd.d);
        break;
    case PLR_LRMTX:
        free(A->dlr.u);
        free(A->dlr.v);
        free(A->dlr.sig);
        break;
    case PLR_HMTX: {
        DPlrMtx *c[4] = {A->dh.nw, A->dh.ne, A->dh.se, A->dh.sw};
        for (int i = 0; i < 4; ++i) {
            dplrclr(c[i]);
            free(c[i]);
        }
        break;
    }
    default:
        break;
    }
    dplrini(A);
}

/* reference src/dplrresizelr.c:15 */
void dplrresizelr(DPlrMtx *A, int capacity)
{
    if (A->type != PLR_LRMTX) {
        fprintf(stderr, "%s: invalid matrix type\n", __func__);
        return;
    }
    if (A->size < 0) {
        fprintf(stderr, "%s: invalid matrix size\n", __func__);
        return;
    }
    if (capacity < 0) {
        fprintf(stderr, "%s: invalid capacity\n", __func__);
        return;
    }
    if (capacity < A->dlr.capacity) return;
    int cap = capacity > 2 * A->dlr.capacity ? capacity : 2 * A->dlr.capacity;
    size_t n = (size_t)cap * (size_t)A->size;
    double *u = malloc((n ? n : 1) * sizeof(double));
    double *v = malloc((n ? n : 1) * sizeof(double));
    double *s = malloc((size_t)(cap ? cap : 1) * sizeof(double));
    if (!u || !v || !s) {
        fprintf(stderr, "%s: memory allocation failed\n", __func__);
        free(u);
        free(v);
        free(s);
        return;
    }
    fill_nan(u, n);
    fill_nan(v, n);
    fill_nan(s, (size_t)cap);
    if (A->dlr.rank > 0) {
        size_t k = (size_t)A->dlr.rank * (size_t)A->size;
        memcpy(u, A->dlr.u, k * sizeof(double));
        memcpy(v, A->dlr.v, k * sizeof(double));
        memcpy(s, A->dlr.sig, (size_t)A->dlr.rank * sizeof(double));
    }
    free(A->dlr.u);
    free(A->dlr.v);
    free(A->dlr.sig);
    A->dlr.capacity = cap;
    A->dlr.u = u;
    A->dlr.v = v;
    A->dlr.sig = s;
}

/* reference src/dplrempty.c:18 */
void dplrempty(DPlrMtx *A, DPlrMtx *B)
{
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "%s: A must be a dense or hierarchical matrix\n", __func__);
        return;
    }
    dplrclr(B);
    if (A->type == PLR_DMTX) {
        dplrinid(A->size, B);
        return;
    }
    dplrinih(A->size, B);
    if (B->type != PLR_HMTX) return;
    dplrinilr(A->size / 2, B->dh.ne, 0);
    dplrinilr(A->size / 2, B->dh.sw, 0);
    dplrempty(A->dh.nw, B->dh.nw);
    dplrempty(A->dh.se, B->dh.se);
}

static void lr_copy(int m, const DPlrMtx *a, DPlrMtx *b)
{
    const int r = a->dlr.rank;
    if (b->dlr.capacity < r) dplrresizelr(b, r);
    if (b->dlr.capacity < r) return; /* allocation failed (reported) */
    if (r > 0) {
        memcpy(b->dlr.sig, a->dlr.sig, (size_t)r * sizeof(double));
        memcpy(b->dlr.u, a->dlr.u, (size_t)r * m * sizeof(double));
        memcpy(b->dlr.v, a->dlr.v, (size_t)r * m * sizeof(double));
    }
    b->dlr.rank = r;
}

/* reference src/dplrcpy.c:19 */
void dplrcpy(int m, DPlrMtx *A, DPlrMtx *B)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size %d\n", __func__, m);
        return;
    }
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "%s: invalid A type\n", __func__);
        return;
    }
    if (A->size != m) {
        fprintf(stderr, "%s: invalid A size\n", __func__);
        return;
    }
    if (A->type == PLR_DMTX) {
        if (B->type != PLR_DMTX || B->size != m) {
            dplrclr(B);
            dplrinid(m, B);
        }
        if (m > 0 && B->dd.d) memcpy(B->dd.d, A->dd.d, (size_t)m * m * sizeof(double));
        return;
    }
    if (B->type != PLR_HMTX || B->size != m) {
        /* new structure; the diagonal blocks are shaped by the recursion */
        dplrclr(B);
        dplrinih(m, B);
        if (B->type != PLR_HMTX) return;
        dplrinilr(m / 2, B->dh.ne, 0);
        dplrinilr(m / 2, B->dh.sw, 0);
    }
    lr_copy(m / 2, A->dh.ne, B->dh.ne);
    lr_copy(m / 2, A->dh.sw, B->dh.sw);
    dplrcpy(m / 2, A->dh.nw, B->dh.nw);
    dplrcpy(m / 2, A->dh.se, B->dh.se);
}

/* ---- dense expansion, reference src/dplrp2d.c:18-143 ------------------- */

static void lr_expand(int m, const DPlrMtx *A, double *B, int ldb)
{
    for (int j = 0; j < m; ++j) memset(B + (size_t)j * ldb, 0, (size_t)m * sizeof(double));
    for (int k = 0; k < A->dlr.rank; ++k) {
        const double *u = A->dlr.u + (size_t)k * m;
        const double *v = A->dlr.v + (size_t)k * m;
        const double s = A->dlr.sig[k];
        for (int j = 0; j < m; ++j) {
            const double t = s * v[j];
            double *col = B + (size_t)j * ldb;
            for (int i = 0; i < m; ++i) col[i] += u[i] * t;
        }
    }
}

void dplrgep2d(int m, DPlrMtx *A, double *B, int ldb)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size %d\n", __func__, m);
        return;
    }
    if (A->size != m) {
        fprintf(stderr, "%s: invalid A size %d\n", __func__, m);
        return;
    }
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "%s: invalid A type\n", __func__);
        return;
    }
    if (ldb < m) {
        fprintf(stderr, "%s: invalid leading dimension\n", __func__);
        return;
    }
    if (A->type == PLR_DMTX) {
        for (int j = 0; j < m; ++j)
            memcpy(B + (size_t)j * ldb, A->dd.d + (size_t)j * m, (size_t)m * sizeof(double));
        return;
    }
    const int h = m / 2;
    lr_expand(h, A->dh.ne, B + (size_t)ldb * h, ldb);
    lr_expand(h, A->dh.sw, B + h, ldb);
    dplrgep2d(h, A->dh.nw, B, ldb);
    dplrgep2d(h, A->dh.se, B + (size_t)ldb * h + h, ldb);
}

/* reference src/dplrrnk.c:22 */
void dplrrnk(int m, DPlrMtx *A, int *ranks)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size %d\n", __func__, m);
        return;
    }
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "%s: invalid A type\n", __func__);
        return;
    }
    if (A->size != m) {
        fprintf(stderr, "%s: invalid A size\n", __func__);
        return;
    }
    if (A->type != PLR_HMTX) return;
    int r = A->dh.ne->dlr.rank > A->dh.sw->dlr.rank ? A->dh.ne->dlr.rank : A->dh.sw->dlr.rank;
    if (*ranks < r) *ranks = r;
    if (A->dh.nw->type != PLR_DMTX && A->dh.se->type != PLR_DMTX) {
        dplrrnk(m / 2, A->dh.nw, ranks + 1);
        dplrrnk(m / 2, A->dh.se, ranks + 1);
    }
}

/* reference src/dplrnz.c:17 */
long int dplrnz(int m, DPlrMtx *A)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size %d\n", __func__, m);
        return -1;
    }
    if (A->size != m) {
        fprintf(stderr, "%s: invalid A size\n", __func__);
        return -1;
    }
    if (A->type != PLR_DMTX && A->type != PLR_HMTX) {
        fprintf(stderr, "%s: invalid A type\n", __func__);
        return -1;
    }
    if (A->type == PLR_DMTX) return (long)m * m;
    const long h = m / 2;
    return 2 * h * (long)A->dh.ne->dlr.rank + 2 * h * (long)A->dh.sw->dlr.rank +
           dplrnz(m / 2, A->dh.nw) + dplrnz(m / 2, A->dh.se);
}
