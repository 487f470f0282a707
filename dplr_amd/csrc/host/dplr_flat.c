/*
 * Tree -> level-wise arrays (layout described in dplr_hip.h).
 *
 * Only uniform trees are accepted: every leaf at the same depth, every
 * internal node with low-rank ne/sw of half its order.  This is the shape
 * dplrd2gep produces (reference src/dplrd2p.c:186-331, exact halving checked
 * at :214).  A non-uniform tree is rejected with a diagnostic.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dplr.h"
#include "../dplr_hip.h"

static int shape_walk(const DPlrMtx *A, int m, int d, int *leafdepth, int *leafsize,
                      int *maxrank, int verbose)
{
    if (A->size != m) {
        if (verbose) fprintf(stderr, "dplr: invalid A size at depth %d\n", d);
        return -1;
    }
    if (A->type == PLR_DMTX) {
        if (m > 0 && !A->dd.d) {
            if (verbose) fprintf(stderr, "dplr: dense leaf without storage\n");
            return -1;
        }
        if (*leafdepth < 0) {
            *leafdepth = d;
            *leafsize = m;
        } else if (*leafdepth != d) {
            if (verbose)
                fprintf(stderr, "dplr: leaves at depths %d and %d; only uniform trees "
                                "are supported on the GPU\n", *leafdepth, d);
            return -2;
        }
        return 0;
    }
    if (A->type != PLR_HMTX) {
        if (verbose) fprintf(stderr, "dplr: invalid A type at depth %d\n", d);
        return -1;
    }
    if (d >= DPLRH_MAXL) {
        if (verbose) fprintf(stderr, "dplr: tree deeper than %d levels\n", DPLRH_MAXL);
        return -2;
    }
    if (m % 2) {
        if (verbose) fprintf(stderr, "dplr: odd order %d at an internal node\n", m);
        return -1;
    }
    const int h = m / 2;
    const DPlrMtx *lr[2] = {A->dh.ne, A->dh.sw};
    for (int i = 0; i < 2; ++i) {
        const DPlrMtx *B = lr[i];
        if (B->type != PLR_LRMTX || B->size != h || B->dlr.rank < 0 ||
            B->dlr.rank > B->dlr.capacity ||
            (B->dlr.rank > 0 && (!B->dlr.u || !B->dlr.v || !B->dlr.sig))) {
            if (verbose) fprintf(stderr, "dplr: invalid low-rank block at depth %d\n", d);
            return -1;
        }
        if (B->dlr.rank > maxrank[d]) maxrank[d] = B->dlr.rank;
    }
    int rc = shape_walk(A->dh.nw, h, d + 1, leafdepth, leafsize, maxrank, verbose);
    if (rc) return rc;
    return shape_walk(A->dh.se, h, d + 1, leafdepth, leafsize, maxrank, verbose);
}

int plr_flat_shape(int m, const void *Av, plr_flat *f, int verbose)
{
    const DPlrMtx *A = (const DPlrMtx *)Av;
    memset(f, 0, sizeof(*f));
    int maxrank[DPLRH_MAXL];
    for (int i = 0; i < DPLRH_MAXL; ++i) maxrank[i] = 0;
    int leafdepth = -1, leafsize = 0;
    int rc = shape_walk(A, m, 0, &leafdepth, &leafsize, maxrank, verbose);
    if (rc) return rc;
    f->N = m;
    f->L = leafdepth;
    f->b = leafsize;
    size_t uo = 0, so = 0;
    for (int d = 0; d < f->L; ++d) {
        int r = maxrank[d];
        f->R[d] = (r + 3) & ~3; /* pad to a multiple of 4 (MFMA K step) */
        f->uoff[d] = uo;
        f->soff[d] = so;
        uo += (size_t)f->R[d] * (size_t)m;
        so += (size_t)f->R[d] << (d + 1);
    }
    f->uoff[f->L] = uo;
    f->soff[f->L] = so;
    return 0;
}

static void fill_lr(const DPlrMtx *B, plr_flat *f, int d, int useg, int vseg)
{
    const int len = B->size, R = f->R[d], r = B->dlr.rank;
    double *U = f->U + f->uoff[d] + (size_t)useg * len;
    double *V = f->V + f->uoff[d] + (size_t)vseg * len;
    double *S = f->S + f->soff[d] + (size_t)useg * R;
    for (int k = 0; k < r; ++k) {
        memcpy(U + (size_t)k * f->N, B->dlr.u + (size_t)k * len, (size_t)len * sizeof(double));
        memcpy(V + (size_t)k * f->N, B->dlr.v + (size_t)k * len, (size_t)len * sizeof(double));
        S[k] = B->dlr.sig[k];
    }
    f->rank[(2 << d) - 2 + useg] = r;
}

/* one block to copy: a leaf (vseg < 0, useg its index) or a low-rank block
 * of level d with the segments of its U and V */
typedef struct
{
    const DPlrMtx *X;
    int d, useg, vseg;
} flat_item;

static size_t list_items(const DPlrMtx *A, int d, int n, flat_item *out, size_t k)
{
    if (A->type == PLR_DMTX) {
        if (out) out[k] = (flat_item){A, d, n, -1};
        return k + 1;
    }
    if (out) {
        out[k] = (flat_item){A->dh.ne, d, 2 * n, 2 * n + 1};
        out[k + 1] = (flat_item){A->dh.sw, d, 2 * n + 1, 2 * n};
    }
    k = list_items(A->dh.nw, d + 1, 2 * n, out, k + 2);
    return list_items(A->dh.se, d + 1, 2 * n + 1, out, k);
}

static void fill_walk(const DPlrMtx *A, plr_flat *f, int d, int n)
{
    if (A->type == PLR_DMTX) {
        const size_t bb = (size_t)f->b * (size_t)f->b;
        if (bb) memcpy(f->D + (size_t)n * bb, A->dd.d, bb * sizeof(double));
        return;
    }
    fill_lr(A->dh.ne, f, d, 2 * n, 2 * n + 1);
    fill_lr(A->dh.sw, f, d, 2 * n + 1, 2 * n);
    fill_walk(A->dh.nw, f, d + 1, 2 * n);
    fill_walk(A->dh.se, f, d + 1, 2 * n + 1);
}

int plr_flat_fill(const void *Av, plr_flat *f)
{
    const size_t nuv = f->uoff[f->L], ns = f->soff[f->L];
    const size_t nd = (size_t)f->N * (size_t)f->b;
    const size_t nr = ((size_t)2 << f->L);
    f->U = calloc(nuv ? nuv : 1, sizeof(double));
    f->V = calloc(nuv ? nuv : 1, sizeof(double));
    f->S = calloc(ns ? ns : 1, sizeof(double));
    f->D = calloc(nd ? nd : 1, sizeof(double));
    f->rank = calloc(nr, sizeof(int));
    if (!f->U || !f->V || !f->S || !f->D || !f->rank) {
        fprintf(stderr, "dplr: host memory allocation failed\n");
        plr_flat_free(f);
        return -3;
    }
    /* the blocks copied in parallel (they fill disjoint parts of the level
     * arrays); the walk alone when the list cannot be allocated */
    const size_t ni = list_items((const DPlrMtx *)Av, 0, 0, NULL, 0);
    flat_item *it = malloc(ni * sizeof(flat_item));
    if (!it) {
        fill_walk((const DPlrMtx *)Av, f, 0, 0);
        return 0;
    }
    list_items((const DPlrMtx *)Av, 0, 0, it, 0);
    const size_t bb = (size_t)f->b * (size_t)f->b;
#pragma omp parallel for schedule(dynamic, 8)
    for (size_t i = 0; i < ni; ++i) {
        if (it[i].vseg < 0) {
            if (bb) memcpy(f->D + (size_t)it[i].useg * bb, it[i].X->dd.d, bb * sizeof(double));
        } else {
            fill_lr(it[i].X, f, it[i].d, it[i].useg, it[i].vseg);
        }
    }
    free(it);
    return 0;
}

void plr_flat_free(plr_flat *f)
{
    free(f->U);
    free(f->V);
    free(f->S);
    free(f->D);
    free(f->rank);
    f->U = f->V = f->S = f->D = NULL;
    f->rank = NULL;
}
