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
