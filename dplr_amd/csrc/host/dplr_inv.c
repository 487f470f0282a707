/*
 * Drop-in dplrgeinv and dplrgeprk: argument checks, the structural reshaping
 * of the output tree and the copy of GPU-computed values into it.
 *
 * The tree shape of B = inv(A) follows the reference's post-order walk
 * (src/dplrinv.c:124-229 and the dplrgeprk appends at :428-429): a node's
 * own Woodbury blocks get A's ranks, then every ancestor's correction
 * appends k = (rank of the ancestor's other off-diagonal block) columns to
 * every low-rank block of the subtree, growing capacity with dplrresizelr
 * exactly when the reference does.  Column order per block is therefore
 * [own | correction from the parent | from the grandparent | ...].
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dplr.h"
#include "../dplr_hip.h"

typedef struct
{
    const plr_flat *f;
    int Rt[DPLRH_MAXL];
    int coff[DPLRH_MAXL][DPLRH_MAXL];
    double *BU[DPLRH_MAXL], *BV[DPLRH_MAXL], *Bsig[DPLRH_MAXL], *BD;
} inv_ctx;

static int rank_at(const plr_flat *f, int d, int s) { return f->rank[(2 << d) - 2 + s]; }

/* dplrgeprk's structural effect: k more columns in every block of the subtree */
/* The output tree is shaped as the reference shapes it (each node created
 * with its own rank, then every ancestor's correction columns appended to
 * its subtrees: at least doubling capacities, src/dplrresizelr.c:54), but
 * on the sizes alone: a node that this sequence would create or grow has
 * its buffers released here and is allocated once, at its final capacity,
 * by alloc_shaped.  The end state is the reference's: the final
 * capacities, every slot below the rank written by fill_inverse, and NaN
 * above it in the buffers this call allocates (a node whose buffers suffice
 * keeps them, slots above the rank untouched). */
static void lr_release(DPlrMtx *X)
{
    free(X->dlr.u);
    free(X->dlr.v);
    free(X->dlr.sig);
    X->dlr.u = X->dlr.v = X->dlr.sig = NULL;
}

/* dplrresizelr on the sizes: the new capacity, the buffers released */
static void lr_grow(DPlrMtx *X, int need)
{
    if (need <= X->dlr.capacity) return;
    X->dlr.capacity = need > 2 * X->dlr.capacity ? need : 2 * X->dlr.capacity;
    lr_release(X);
}

static void grow_subtree(DPlrMtx *X, int k)
{
    if (k == 0 || X->type != PLR_HMTX) return;
    DPlrMtx *lr[2] = {X->dh.ne, X->dh.sw};
    for (int i = 0; i < 2; ++i) {
        lr_grow(lr[i], lr[i]->dlr.rank + k);
        lr[i]->dlr.rank += k;
    }
    grow_subtree(X->dh.nw, k);
    grow_subtree(X->dh.se, k);
}

static void shape_inverse(const DPlrMtx *A, DPlrMtx *B, int m, int d)
{
    if (A->type == PLR_DMTX) {
        if (B->type != PLR_DMTX || B->size != m) {
            dplrclr(B);
            dplrinid(m, B);
        }
        return;
    }
    if (B->type != PLR_HMTX || B->size != m) {
        dplrclr(B);
        dplrinih(m, B);
    }
    shape_inverse(A->dh.nw, B->dh.nw, m / 2, d + 1);
    shape_inverse(A->dh.se, B->dh.se, m / 2, d + 1);
    DPlrMtx *Bx[2] = {B->dh.ne, B->dh.sw};
    const DPlrMtx *Ax[2] = {A->dh.ne, A->dh.sw};
    for (int i = 0; i < 2; ++i) {
        if (Bx[i]->type != PLR_LRMTX || Bx[i]->size != m / 2) {
            /* dplrinilr on the sizes */
            dplrclr(Bx[i]);
            Bx[i]->type = PLR_LRMTX;
            Bx[i]->size = m / 2;
            Bx[i]->dlr.rank = 0;
            Bx[i]->dlr.capacity = Ax[i]->dlr.rank;
        }
        lr_grow(Bx[i], Ax[i]->dlr.rank);
    }
    const int r1 = A->dh.ne->dlr.rank, r2 = A->dh.sw->dlr.rank;
    if (r1 == 0 && r2 == 0) return;
    B->dh.ne->dlr.rank = r1;
    B->dh.sw->dlr.rank = r2;
    grow_subtree(B->dh.nw, r2);
    grow_subtree(B->dh.se, r1);
}

/* Copy the level-d' block whose U lives on segment q: own columns, then the
 * corrections from levels d'-1 .. 0. */
static void fill_block(DPlrMtx *X, const inv_ctx *c, int dp, int q)
{
    const plr_flat *f = c->f;
    const int len = X->size, Rt = c->Rt[dp];
    const double *bu = c->BU[dp] + (size_t)q * Rt * len;
    const double *bv = c->BV[dp] + (size_t)q * Rt * len;
    const double *bs = c->Bsig[dp] + (size_t)q * Rt;
    int pos = 0;
    const int rown = rank_at(f, dp, q);
    for (int j = 0; j < rown; ++j, ++pos) {
        memcpy(X->dlr.u + (size_t)pos * len, bu + (size_t)j * len, len * sizeof(double));
        memcpy(X->dlr.v + (size_t)pos * len, bv + (size_t)j * len, len * sizeof(double));
        X->dlr.sig[pos] = bs[j];
    }
    for (int d = dp - 1; d >= 0; --d) {
        const int s = q >> (dp - d);
        const int k = rank_at(f, d, s ^ 1);
        for (int j = 0; j < k; ++j, ++pos) {
            const int col = c->coff[dp][d] + j;
            memcpy(X->dlr.u + (size_t)pos * len, bu + (size_t)col * len, len * sizeof(double));
            memcpy(X->dlr.v + (size_t)pos * len, bv + (size_t)col * len, len * sizeof(double));
            X->dlr.sig[pos] = bs[col];
        }
    }
    if (pos != X->dlr.rank)
        fprintf(stderr, "dplrgeinv: internal rank mismatch (%d vs %d)\n", pos, X->dlr.rank);
}

/* The output's blocks in a flat list: leaves (level -1) and low-rank
 * blocks with their level and segment, for the parallel passes below. */
typedef struct
{
    DPlrMtx *X;
    int d, q;
} inv_blk;

static size_t list_blocks(DPlrMtx *B, int d, int n, inv_blk *out, size_t k)
{
    if (B->type == PLR_DMTX) {
        if (out) out[k] = (inv_blk){B, -1, n};
        return k + 1;
    }
    if (B->type != PLR_HMTX) return k;
    if (out) {
        out[k] = (inv_blk){B->dh.ne, d, 2 * n};
        out[k + 1] = (inv_blk){B->dh.sw, d, 2 * n + 1};
    }
    k = list_blocks(B->dh.nw, d + 1, 2 * n, out, k + 2);
    return list_blocks(B->dh.se, d + 1, 2 * n + 1, out, k);
}

static void nan_tail(double *p, size_t from, size_t to)
{
    for (size_t i = from; i < to; ++i) p[i] = NAN;
}

/* Buffers of the low-rank blocks the shaping released or created, at their
 * final capacity; NaN above the rank (below it fill_inverse writes). */
static int alloc_shaped(inv_blk *bl, size_t nb)
{
    int fail = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(| : fail)
    for (size_t i = 0; i < nb; ++i) {
        DPlrMtx *X = bl[i].X;
        if (bl[i].d < 0 || X->dlr.u || X->dlr.capacity == 0) continue;
        const size_t len = (size_t)X->size, cap = (size_t)X->dlr.capacity, r = (size_t)X->dlr.rank;
        X->dlr.u = malloc(len * cap * sizeof(double));
        X->dlr.v = malloc(len * cap * sizeof(double));
        X->dlr.sig = malloc(cap * sizeof(double));
        if (!X->dlr.u || !X->dlr.v || !X->dlr.sig) {
            lr_release(X);
            fail = 1;
            continue;
        }
        nan_tail(X->dlr.u, r * len, cap * len);
        nan_tail(X->dlr.v, r * len, cap * len);
        nan_tail(X->dlr.sig, r, cap);
    }
    return fail;
}

static void fill_inverse(inv_blk *bl, size_t nb, const inv_ctx *c)
{
    const size_t bb = (size_t)c->f->b * c->f->b;
#pragma omp parallel for schedule(dynamic, 8)
    for (size_t i = 0; i < nb; ++i) {
        if (bl[i].d < 0) {
            if (bb) memcpy(bl[i].X->dd.d, c->BD + (size_t)bl[i].q * bb, bb * sizeof(double));
        } else {
            fill_block(bl[i].X, c, bl[i].d, bl[i].q);
        }
    }
}

static void nan_fill(DPlrMtx *X)
{
    if (X->type == PLR_DMTX) {
        for (size_t i = 0; i < (size_t)X->size * X->size; ++i) X->dd.d[i] = NAN;
        return;
    }
    DPlrMtx *lr[2] = {X->dh.ne, X->dh.sw};
    for (int i = 0; i < 2; ++i) {
        const size_t n = (size_t)lr[i]->size * lr[i]->dlr.rank;
        for (size_t j = 0; j < n; ++j) lr[i]->dlr.u[j] = lr[i]->dlr.v[j] = NAN;
        for (int j = 0; j < lr[i]->dlr.rank; ++j) lr[i]->dlr.sig[j] = NAN;
    }
    nan_fill(X->dh.nw);
    nan_fill(X->dh.se);
}

/* With a zero threshold the reference normalises a zero column by dividing
 * by its zero norm (src/dplrinv.c:333-348, :378-404), and the NaN it leaves
 * in a subtree reaches every later Woodbury step that reads that subtree
 * (src/dplrinv.c:245-254 apply the subtree inverses to U and V): the whole
 * subtree of the shallowest ancestor that performs one becomes NaN.  The
 * GPU values already hold the NaN of the flagged node itself. */
static void nan_spread(DPlrMtx *B, const plr_flat *f, const int *flag)
{
    for (int d = 1; d < f->L; ++d)
        for (int n = 0; n < (1 << d); ++n) {
            if (!flag[(1 << d) - 1 + n]) continue;
            for (int da = 0; da < d; ++da) {
                const int na = n >> (d - da);
                if (rank_at(f, da, 2 * na) + rank_at(f, da, 2 * na + 1) == 0) continue;
                DPlrMtx *X = B;
                for (int k = da - 1; k >= 0; --k) X = ((na >> k) & 1) ? X->dh.se : X->dh.nw;
                nan_fill(X);
                break;
            }
        }
}

static void touch_pages(double *p, size_t n)
{
    const size_t step = 4096 / sizeof(double);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i += step) p[i] = 0.0;
}

/* reference src/dplrinv.c:93 */
void dplrgeinv(int m, DPlrMtx *A, DPlrMtx *B, double eps, int *info)
{
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size: %d\n", __func__, m);
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
    *info = -1;
    plr_flat f;
    if (plr_flat_shape(m, A, &f, 1) || plr_flat_fill(A, &f)) return;
    dplrh_plan *p = NULL;
    inv_ctx c;
    memset(&c, 0, sizeof(c));
    c.f = &f;
    int *nanflag = NULL;
    int rc = dplrh_plan_create(&f, &p), finfo = 0;
    if (!rc && m > 0) rc = dplrh_factor(p, &finfo, NULL);
    if (rc) {
        fprintf(stderr, "%s: GPU execution failed (%d): %s\n", __func__, rc, dplrh_last_error());
        goto out;
    }
    if (finfo) {
        fprintf(stderr, "%s: failed to invert a dense matrix: %d\n", __func__, finfo);
        *info = finfo;
        goto out;
    }
    for (int dp = 0; dp < f.L; ++dp) {
        int acc = f.R[dp];
        for (int d = dp - 1; d >= 0; --d) {
            c.coff[dp][d] = acc;
            acc += f.R[d];
        }
        c.Rt[dp] = acc;
        const size_t nseg = (size_t)2 << dp;
        c.BU[dp] = malloc(((size_t)f.N * acc + 1) * sizeof(double));
        c.BV[dp] = malloc(((size_t)f.N * acc + 1) * sizeof(double));
        c.Bsig[dp] = malloc((nseg * acc + 1) * sizeof(double));
        if (!c.BU[dp] || !c.BV[dp] || !c.Bsig[dp]) {
            fprintf(stderr, "%s: host memory allocation failed\n", __func__);
            goto out;
        }
    }
    c.BD = malloc(((size_t)f.N * f.b + 1) * sizeof(double));
    nanflag = calloc((size_t)1 << f.L, sizeof(int));
    if (!c.BD || !nanflag) {
        fprintf(stderr, "%s: host memory allocation failed\n", __func__);
        goto out;
    }
    /* the pages of the staging buffers mapped by all threads before the
     * download writes them (a first touch inside the copy is one thread's) */
    for (int dp = 0; dp < f.L; ++dp) {
        touch_pages(c.BU[dp], (size_t)f.N * c.Rt[dp]);
        touch_pages(c.BV[dp], (size_t)f.N * c.Rt[dp]);
    }
    touch_pages(c.BD, (size_t)f.N * f.b);
    if (m > 0) {
        rc = dplrh_geinv(p, eps, c.Rt, c.coff, c.BU, c.BV, c.Bsig, c.BD, nanflag, NULL);
        if (rc) {
            fprintf(stderr, "%s: GPU execution failed (%d): %s\n", __func__, rc, dplrh_last_error());
            goto out;
        }
    }
    shape_inverse(A, B, m, 0);
    {
        const size_t nb = list_blocks(B, 0, 0, NULL, 0);
        inv_blk *bl = malloc((nb ? nb : 1) * sizeof(inv_blk));
        if (!bl) {
            fprintf(stderr, "%s: host memory allocation failed\n", __func__);
            goto out;
        }
        list_blocks(B, 0, 0, bl, 0);
        if (alloc_shaped(bl, nb)) {
            fprintf(stderr, "%s: memory allocation failed\n", __func__);
            free(bl);
            goto out;
        }
        if (m > 0) {
            fill_inverse(bl, nb, &c);
            nan_spread(B, &f, nanflag);
        }
        free(bl);
    }
    *info = 0;
out:
    for (int d = 0; d < DPLRH_MAXL; ++d) {
        free(c.BU[d]);
        free(c.BV[d]);
        free(c.Bsig[d]);
    }
    free(c.BD);
    free(nanflag);
    dplrh_plan_destroy(p);
    plr_flat_free(&f);
}

/* ---- rank-k update ------------------------------------------------------ */

static void gather_leaves(const DPlrMtx *A, int b, int n, double *D)
{
    if (A->type == PLR_DMTX) {
        if (b) memcpy(D + (size_t)n * b * b, A->dd.d, (size_t)b * b * sizeof(double));
        return;
    }
    gather_leaves(A->dh.nw, b, 2 * n, D);
    gather_leaves(A->dh.se, b, 2 * n + 1, D);
}

static void scatter_leaves(DPlrMtx *A, int b, int n, const double *D)
{
    if (A->type == PLR_DMTX) {
        if (b) memcpy(A->dd.d, D + (size_t)n * b * b, (size_t)b * b * sizeof(double));
        return;
    }
    scatter_leaves(A->dh.nw, b, 2 * n, D);
    scatter_leaves(A->dh.se, b, 2 * n + 1, D);
}

typedef struct
{
    int k;
    double **Un, **Vn, **sig;
} prk_ctx;

static void append_block(DPlrMtx *X, const prk_ctx *c, int d, int q)
{
    const int len = X->size, r0 = X->dlr.rank, k = c->k;
    if (X->dlr.capacity < r0 + k) dplrresizelr(X, r0 + k);
    const double *bu = c->Un[d] + (size_t)q * k * len;
    const double *bv = c->Vn[d] + (size_t)q * k * len;
    memcpy(X->dlr.u + (size_t)r0 * len, bu, (size_t)k * len * sizeof(double));
    memcpy(X->dlr.v + (size_t)r0 * len, bv, (size_t)k * len * sizeof(double));
    memcpy(X->dlr.sig + r0, c->sig[d] + (size_t)q * k, (size_t)k * sizeof(double));
    X->dlr.rank = r0 + k;
}

static void append_walk(DPlrMtx *A, const prk_ctx *c, int d, int n)
{
    if (A->type != PLR_HMTX) return;
    append_block(A->dh.ne, c, d, 2 * n);
    append_block(A->dh.sw, c, d, 2 * n + 1);
    append_walk(A->dh.nw, c, d + 1, 2 * n);
    append_walk(A->dh.se, c, d + 1, 2 * n + 1);
}

/* Host buffers of dplrgeprk, kept from call to call while the thread's
 * pool is on (as the device ones, dplrh_prk_pool): the pages a call's
 * downloads write are then mapped already, not faulted in afresh. */
enum { PRK_NBUF = 4 + 3 * DPLRH_MAXL };
typedef struct
{
    int on;
    double *p[PRK_NBUF];
    size_t cap[PRK_NBUF];
} prk_hpool;
static _Thread_local prk_hpool g_hp;

static double *prk_alloc(int i, size_t n)
{
    if (!g_hp.on) return malloc((n ? n : 1) * sizeof(double));
    if (g_hp.cap[i] < n) {
        free(g_hp.p[i]);
        g_hp.p[i] = malloc((n ? n : 1) * sizeof(double));
        g_hp.cap[i] = g_hp.p[i] ? n : 0;
    }
    return g_hp.p[i];
}

static void prk_release(double *p)
{
    if (!g_hp.on) free(p);
}

void plr_prk_pool(int on)
{
    if (!on)
        for (int i = 0; i < PRK_NBUF; ++i) {
            free(g_hp.p[i]);
            g_hp.p[i] = NULL;
            g_hp.cap[i] = 0;
        }
    g_hp.on = on != 0;
    dplrh_prk_pool(on);
}

/* reference src/dplrprk.c:21 */
void dplrgeprk(int m, int k, double *alpha, double *u, int ldu, double *v, int ldv, double *s,
               int incs, DPlrMtx *A, double eps)
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
    if (ldu < m || ldv < m) {
        fprintf(stderr, "%s: invalid leading dimensions\n", __func__);
        return;
    }
    if (*alpha == 0.0 || k == 0) return;
    if (k < 0) {
        fprintf(stderr, "%s: invalid rank %d\n", __func__, k);
        return;
    }
    plr_flat f;
    if (plr_flat_shape(m, A, &f, 1)) return;
    const int N = f.N, L = f.L, b = f.b;
    if (N == 0) return;
    const size_t nk = (size_t)N * k;
    double *hu = prk_alloc(0, nk), *hv = prk_alloc(1, nk);
    double *hs = prk_alloc(2, (size_t)k);
    double *hD = prk_alloc(3, (size_t)N * b);
    double *Un[DPLRH_MAXL] = {0}, *Vn[DPLRH_MAXL] = {0}, *sig[DPLRH_MAXL] = {0};
    int ok = hu && hv && hs && hD;
    for (int d = 0; d < L && ok; ++d) {
        Un[d] = prk_alloc(4 + 3 * d, nk);
        Vn[d] = prk_alloc(5 + 3 * d, nk);
        sig[d] = prk_alloc(6 + 3 * d, ((size_t)2 << d) * k);
        ok = Un[d] && Vn[d] && sig[d];
    }
    if (!ok) {
        fprintf(stderr, "%s: host memory allocation failed\n", __func__);
    } else {
        for (int j = 0; j < k; ++j) {
            memcpy(hu + (size_t)j * N, u + (size_t)j * ldu, (size_t)N * sizeof(double));
            memcpy(hv + (size_t)j * N, v + (size_t)j * ldv, (size_t)N * sizeof(double));
            hs[j] = s[(size_t)j * incs];
        }
        gather_leaves(A, b, 0, hD);
        int rc = dplrh_geprk(N, L, b, k, *alpha, eps, hu, hv, hs, hD, Un, Vn, sig, NULL);
        if (rc) {
            fprintf(stderr, "%s: GPU execution failed (%d): %s\n", __func__, rc, dplrh_last_error());
        } else {
            prk_ctx c = {k, Un, Vn, sig};
            /* the blocks take their columns in parallel (each writes its own
             * buffers); the walk alone when the list cannot be allocated */
            const size_t nb = list_blocks(A, 0, 0, NULL, 0);
            inv_blk *bl = malloc((nb ? nb : 1) * sizeof(inv_blk));
            if (bl) {
                list_blocks(A, 0, 0, bl, 0);
#pragma omp parallel for schedule(dynamic, 8)
                for (size_t i = 0; i < nb; ++i)
                    if (bl[i].d >= 0) append_block(bl[i].X, &c, bl[i].d, bl[i].q);
                free(bl);
            } else {
                append_walk(A, &c, 0, 0);
            }
            scatter_leaves(A, b, 0, hD);
        }
    }
    for (int d = 0; d < DPLRH_MAXL; ++d) {
        prk_release(Un[d]);
        prk_release(Vn[d]);
        prk_release(sig[d]);
    }
    prk_release(hu);
    prk_release(hv);
    prk_release(hs);
    prk_release(hD);
}
