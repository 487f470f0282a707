/*
 * Deterministic synthetic inputs.
 *
 * Random numbers come from a counter-based generator: the value for
 * (seed, stream, block, index) is a 64-bit mix of those integers, so any
 * block can be generated independently (and in parallel) with the same
 * result.  Normals use Box-Muller on two uniforms.  Orthonormal factors are
 * the thin Q of a Householder QR of an N(0,1) matrix; this file is compiled
 * with -ffp-contract=off so the result does not depend on FMA contraction.
 *
 * Block ids: the node with index n at depth d has id 2^d - 1 + n; its ne.U,
 * ne.V, sw.U and sw.V use parts 0..3, a leaf uses part 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dplr.h"
#include "dplrx.h"

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t gen_key(uint64_t seed, uint32_t stream, uint64_t block)
{
    return mix64(mix64(seed) ^ ((uint64_t)stream << 48) ^ mix64(block + 0x5851f42d4c957f2dULL));
}

static inline double unif(uint64_t key, uint64_t ctr)
{
    const uint64_t z = mix64(key ^ mix64(ctr));
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0); /* (0,1) */
}

static inline double gauss(uint64_t key, uint64_t idx)
{
    const double u1 = unif(key, 2 * idx), u2 = unif(key, 2 * idx + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Thin Q (n x r, column-major) of the Householder QR of G (n x r); G is
 * overwritten.  Requires r <= n. */
static void householder_q(int n, int r, double *G, double *Q, double *tau)
{
    for (int j = 0; j < r; ++j) {
        double *x = G + (size_t)j * n + j;
        const int len = n - j;
        double nrm = 0.0;
        for (int i = 0; i < len; ++i) nrm += x[i] * x[i];
        nrm = sqrt(nrm);
        if (nrm == 0.0) {
            tau[j] = 0.0;
            continue;
        }
        const double alpha = x[0];
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        const double scal = 1.0 / (alpha - beta);
        for (int i = 1; i < len; ++i) x[i] *= scal;
        tau[j] = (beta - alpha) / beta;
        x[0] = 1.0; /* v = [1; x(1:)] while it is applied */
        for (int c = j + 1; c < r; ++c) {
            double *y = G + (size_t)c * n + j;
            double dot = 0.0;
            for (int i = 0; i < len; ++i) dot += x[i] * y[i];
            dot *= tau[j];
            for (int i = 0; i < len; ++i) y[i] -= dot * x[i];
        }
        x[0] = beta;
    }
    for (int c = 0; c < r; ++c) {
        double *q = Q + (size_t)c * n;
        memset(q, 0, (size_t)n * sizeof(double));
        q[c] = 1.0;
    }
    for (int j = r - 1; j >= 0; --j) {
        if (tau[j] == 0.0) continue;
        double *v = G + (size_t)j * n + j;
        const double keep = v[0];
        v[0] = 1.0;
        const int len = n - j;
        for (int c = j; c < r; ++c) {
            double *y = Q + (size_t)c * n + j;
            double dot = 0.0;
            for (int i = 0; i < len; ++i) dot += v[i] * y[i];
            dot *= tau[j];
            for (int i = 0; i < len; ++i) y[i] -= dot * v[i];
        }
        v[0] = keep;
    }
}

typedef struct
{
    double *dst; /* factor (len x r) or leaf (len x len) */
    int len, r;
    uint64_t block;
    int leaf;
} gen_task;

typedef struct
{
    gen_task *t;
    size_t n, cap;
} task_list;

static int push(task_list *l, gen_task t)
{
    if (l->n == l->cap) {
        size_t cap = l->cap ? 2 * l->cap : 256;
        gen_task *p = realloc(l->t, cap * sizeof(gen_task));
        if (!p) return -1;
        l->t = p;
        l->cap = cap;
    }
    l->t[l->n++] = t;
    return 0;
}

static int build(DPlrMtx *A, int m, int d, int L, long n, int rank, task_list *tl)
{
    const uint64_t id = ((uint64_t)1 << d) - 1 + (uint64_t)n;
    if (d == L) {
        dplrinid(m, A);
        if (m > 0 && !A->dd.d) return -1;
        gen_task t = {A->dd.d, m, m, id << 2, 1};
        return push(tl, t);
    }
    const int h = m / 2;
    const int r = rank < h ? rank : h;
    dplrinih(m, A);
    if (A->type != PLR_HMTX) return -1;
    dplrinilr(h, A->dh.ne, r);
    dplrinilr(h, A->dh.sw, r);
    A->dh.ne->dlr.rank = r;
    A->dh.sw->dlr.rank = r;
    for (int j = 0; j < r; ++j) {
        A->dh.ne->dlr.sig[j] = pow(10.0, -j / 4.0);
        A->dh.sw->dlr.sig[j] = pow(10.0, -j / 4.0);
    }
    if (r > 0) {
        gen_task t0 = {A->dh.ne->dlr.u, h, r, (id << 2) | 0, 0};
        gen_task t1 = {A->dh.ne->dlr.v, h, r, (id << 2) | 1, 0};
        gen_task t2 = {A->dh.sw->dlr.u, h, r, (id << 2) | 2, 0};
        gen_task t3 = {A->dh.sw->dlr.v, h, r, (id << 2) | 3, 0};
        if (push(tl, t0) || push(tl, t1) || push(tl, t2) || push(tl, t3)) return -1;
    }
    if (build(A->dh.nw, h, d + 1, L, 2 * n, rank, tl)) return -1;
    return build(A->dh.se, h, d + 1, L, 2 * n + 1, rank, tl);
}

int dplrx_gen_random(int m, int nlvl, int rank, double shift, unsigned long long seed,
                     DPlrMtx *A)
{
    if (m < 0 || nlvl < 0 || nlvl > 30 || rank < 0 || (m % (1 << nlvl)) != 0) {
        fprintf(stderr, "%s: invalid arguments (m=%d nlvl=%d rank=%d)\n", __func__, m, nlvl,
                rank);
        return -1;
    }
    dplrclr(A);
    task_list tl = {NULL, 0, 0};
    if (build(A, m, 0, nlvl, 0, rank, &tl)) {
        fprintf(stderr, "%s: allocation failed\n", __func__);
        free(tl.t);
        return -3;
    }
    int fail = 0;
    const uint64_t key0 = (uint64_t)seed;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : fail)
    for (long i = 0; i < (long)tl.n; ++i) {
        const gen_task *t = &tl.t[i];
        const uint64_t key = gen_key(key0, 0, t->block);
        if (t->leaf) {
            const int b = t->len;
            const double sc = 1.0 / sqrt((double)b);
            for (size_t e = 0; e < (size_t)b * b; ++e) t->dst[e] = gauss(key, e) * sc;
            for (int j = 0; j < b; ++j) t->dst[(size_t)j * b + j] += shift;
        } else {
            const size_t ne = (size_t)t->len * t->r;
            double *G = malloc(ne * sizeof(double));
            double *tau = malloc((size_t)t->r * sizeof(double));
            if (!G || !tau) {
                fail |= 1;
            } else {
                for (size_t e = 0; e < ne; ++e) G[e] = gauss(key, e);
                householder_q(t->len, t->r, G, t->dst, tau);
            }
            free(G);
            free(tau);
        }
    }
    free(tl.t);
    if (fail) {
        fprintf(stderr, "%s: allocation failed\n", __func__);
        return -3;
    }
    return 0;
}

int dplrx_gen_normal(size_t n, unsigned long long seed, unsigned int stream, double *out)
{
    const uint64_t key = gen_key((uint64_t)seed, stream, 0);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) out[i] = gauss(key, (uint64_t)i);
    return 0;
}
