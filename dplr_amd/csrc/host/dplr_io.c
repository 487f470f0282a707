/*
 * PLR file format: a DPlrMtx tree on disk (the reference keeps trees in
 * memory only).  Little-endian, 8-byte aligned records:
 *
 *   header   "DPLRTREE" (8 bytes), uint32 version (1), uint32 0, int64 order
 *   node     int32 type (1 hierarchical, 2 dense, 3 low-rank), int32 size,
 *            then by type
 *              1: the children nw, ne, se, sw (pre-order)
 *              2: size*size doubles, column-major
 *              3: int32 rank, int32 0, rank*size doubles of u (column-major),
 *                 the same of v, rank doubles of sig
 *
 * Only the valid columns of a low-rank block are stored; a loaded block has
 * capacity == rank.  Loading replaces the target tree.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dplr.h"
#include "dplrx.h"

static const char kMagic[8] = {'D', 'P', 'L', 'R', 'T', 'R', 'E', 'E'};

static int put(FILE *f, const void *p, size_t bytes)
{
    return bytes == 0 || fwrite(p, 1, bytes, f) == bytes ? 0 : -2;
}

static int get(FILE *f, void *p, size_t bytes)
{
    return bytes == 0 || fread(p, 1, bytes, f) == bytes ? 0 : -2;
}

static int save_node(FILE *f, const DPlrMtx *A)
{
    int32_t hdr[2] = {(int32_t)A->type, A->size};
    int rc = put(f, hdr, sizeof(hdr));
    if (rc) return rc;
    const size_t m = (size_t)(A->size > 0 ? A->size : 0);
    switch (A->type) {
    case PLR_HMTX: {
        const DPlrMtx *c[4] = {A->dh.nw, A->dh.ne, A->dh.se, A->dh.sw};
        for (int i = 0; i < 4 && !rc; ++i) rc = c[i] ? save_node(f, c[i]) : -4;
        return rc;
    }
    case PLR_DMTX:
        return m ? put(f, A->dd.d, m * m * sizeof(double)) : 0;
    case PLR_LRMTX: {
        const int32_t rk[2] = {A->dlr.rank, 0};
        const size_t r = (size_t)(A->dlr.rank > 0 ? A->dlr.rank : 0);
        if ((rc = put(f, rk, sizeof(rk)))) return rc;
        if (r == 0) return 0;
        if ((rc = put(f, A->dlr.u, m * r * sizeof(double)))) return rc;
        if ((rc = put(f, A->dlr.v, m * r * sizeof(double)))) return rc;
        return put(f, A->dlr.sig, r * sizeof(double));
    }
    default:
        return -4; /* an unpopulated node has no file form */
    }
}

int dplrx_save(const char *path, int m, DPlrMtx *A)
{
    if (m < 0 || A->size != m || (A->type != PLR_HMTX && A->type != PLR_DMTX)) {
        fprintf(stderr, "%s: invalid matrix\n", __func__);
        return -1;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "%s: cannot open %s\n", __func__, path);
        return -2;
    }
    const uint32_t ver[2] = {1, 0};
    const int64_t order = m;
    int rc = put(f, kMagic, sizeof(kMagic));
    if (!rc) rc = put(f, ver, sizeof(ver));
    if (!rc) rc = put(f, &order, sizeof(order));
    if (!rc) rc = save_node(f, A);
    if (fclose(f) != 0 && !rc) rc = -2;
    if (rc) fprintf(stderr, "%s: writing %s failed (%d)\n", __func__, path, rc);
    return rc;
}

/* sizes are checked against the parent's before any allocation */
static int load_node(FILE *f, DPlrMtx *A, int expect, int depth)
{
    int32_t hdr[2];
    int rc = get(f, hdr, sizeof(hdr));
    if (rc) return rc;
    const int type = hdr[0], m = hdr[1];
    if (m != expect || m < 0 || depth > 40) return -3;
    switch (type) {
    case PLR_HMTX:
        if (m % 2 != 0) return -3;
        dplrinih(m, A);
        if (A->type != PLR_HMTX) return -5;
        {
            DPlrMtx *c[4] = {A->dh.nw, A->dh.ne, A->dh.se, A->dh.sw};
            for (int i = 0; i < 4 && !rc; ++i) rc = load_node(f, c[i], m / 2, depth + 1);
        }
        return rc;
    case PLR_DMTX:
        dplrinid(m, A);
        if (m && !A->dd.d) return -5;
        return m ? get(f, A->dd.d, (size_t)m * m * sizeof(double)) : 0;
    case PLR_LRMTX: {
        int32_t rk[2];
        if ((rc = get(f, rk, sizeof(rk)))) return rc;
        const int r = rk[0];
        if (r < 0 || (m > 0 && r > m) || (m == 0 && r != 0)) return -3;
        dplrinilr(m, A, r);
        if (r && (!A->dlr.u || !A->dlr.v || !A->dlr.sig)) return -5;
        A->dlr.rank = r;
        if (r == 0) return 0;
        if ((rc = get(f, A->dlr.u, (size_t)m * r * sizeof(double)))) return rc;
        if ((rc = get(f, A->dlr.v, (size_t)m * r * sizeof(double)))) return rc;
        return get(f, A->dlr.sig, (size_t)r * sizeof(double));
    }
    default:
        return -3;
    }
}

int dplrx_load(const char *path, DPlrMtx *A, int *m)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "%s: cannot open %s\n", __func__, path);
        return -2;
    }
    char magic[8];
    uint32_t ver[2];
    int64_t order = -1;
    int rc = get(f, magic, sizeof(magic));
    if (!rc) rc = get(f, ver, sizeof(ver));
    if (!rc) rc = get(f, &order, sizeof(order));
    if (!rc && (memcmp(magic, kMagic, sizeof(kMagic)) || ver[0] != 1 || order < 0 ||
                order > 0x7fffffff))
        rc = -3;
    dplrclr(A);
    if (!rc) rc = load_node(f, A, (int)order, 0);
    if (!rc) {
        char extra;
        if (fread(&extra, 1, 1, f) == 1) rc = -3; /* trailing bytes: not this format */
    }
    fclose(f);
    if (rc) {
        dplrclr(A);
        fprintf(stderr, "%s: %s is not a readable PLR file (%d)\n", __func__, path, rc);
        return rc;
    }
    if (m) *m = (int)order;
    return 0;
}
