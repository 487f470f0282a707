/*
 * dplr.h -- drop-in C API of the MI355X PLR (HODLR) library.
 *
 * A PLR matrix of order m is a perfect binary tree.  An internal node is a
 * 2x2 block matrix [nw ne; sw se] whose off-diagonal blocks ne and sw are
 * low rank, U*diag(sig)*V^T, and whose diagonal blocks recurse; a leaf is a
 * dense column-major block.  Every split halves the order exactly.
 *
 * The node layout, the function names, their argument order and meaning,
 * the pass-by-pointer scalars, the column-major storage, the stderr
 * diagnostics and the early-return-on-error behaviour are those of the
 * reference header (include/dplr.h and include/dplr_internal.h of dplr), so
 * that code written against it links against this library unchanged.  All
 * arguments are host memory.  The arithmetic of dplrgepmv, dplrgepdmm,
 * dplrgeinv and dplrgeprk runs on the GPU (gfx950) through the HIP layer of
 * this library; the device-resident interface for callers that keep data in
 * HBM is declared in dplrx.h.
 */
#ifndef DPLR_AMD_DPLR_H
#define DPLR_AMD_DPLR_H

#ifdef __cplusplus
extern "C" {
#endif

/* Node tag.  Values follow include/dplr_internal.h:32-38 of the reference. */
typedef enum
{
    PLR_UMTX = 0,   /* not yet populated */
    PLR_HMTX,       /* 2x2 hierarchical node */
    PLR_DMTX,       /* dense leaf */
    PLR_LRMTX       /* low-rank block */
} PlrMtxType;

typedef struct DPlrMtxStruct DPlrMtx;

/* Children of a hierarchical node (reference dplr_internal.h:48-51).  The
 * field order nw, ne, se, sw is part of the ABI. */
typedef struct
{
    DPlrMtx *nw, *ne, *se, *sw;
} DPlrHMtx;

/* Dense leaf: size x size, column-major (reference dplr_internal.h:56-59). */
typedef struct
{
    double *d;
} DPlrDMtx;

/* Low-rank block U*diag(sig)*V^T (reference dplr_internal.h:64-70).  u and v
 * are size x capacity column-major with leading dimension size; columns
 * [0, rank) are valid. */
typedef struct
{
    int rank;
    int capacity;
    double *u, *v;
    double *sig;
} DPlrLRMtx;

/* Generic node (reference dplr_internal.h:75-87): 40 bytes on x86-64. */
struct DPlrMtxStruct
{
    PlrMtxType type;
    int size;
    union
    {
        DPlrHMtx dh;
        DPlrDMtx dd;
        DPlrLRMtx dlr;
    };
};

/* ---- lifecycle (host memory; reference src/dplrini.c, src/dplrresizelr.c) */

/* Replaces dplrini, reference src/dplrini.c:18.  Zero the node, type UMTX. */
void dplrini(DPlrMtx *A);

/* Replaces dplrclr, reference src/dplrini.c:183.  Frees every buffer of the
 * tree.  Unlike the reference it also frees the child nodes it allocated and
 * resets A to an empty UMTX node, so a cleared node can be re-shaped. */
void dplrclr(DPlrMtx *A);

/* Replaces dplrinid, reference src/dplrini.c:35.  Dense m x m leaf, buffer
 * filled with NaN. */
void dplrinid(int m, DPlrMtx *A);

/* Replaces dplrinilr, reference src/dplrini.c:80.  Low-rank node of order m,
 * rank 0, given capacity, buffers filled with NaN. */
void dplrinilr(int m, DPlrMtx *A, int capacity);

/* Replaces dplrinih, reference src/dplrini.c:142.  Hierarchical node of even
 * order m with four empty (UMTX) children. */
void dplrinih(int m, DPlrMtx *A);

/* Replaces dplrresizelr, reference src/dplrresizelr.c:15.  Grows capacity
 * to max(capacity, 2*old) unless capacity < old; keeps the first rank
 * columns and NaN-fills the rest. */
void dplrresizelr(DPlrMtx *A, int capacity);

/* Replaces dplrempty, reference src/dplrempty.c:18.  B becomes a matrix of
 * A's structure (sizes and node types) with rank-0 low-rank blocks and
 * NaN-filled leaves; B's previous contents are freed. */
void dplrempty(DPlrMtx *A, DPlrMtx *B);

/* Replaces dplrcpy, reference src/dplrcpy.c:19.  B := A for an m x m A;
 * B (initialised) is reshaped where its structure differs and its
 * low-rank blocks grow to A's ranks. */
void dplrcpy(int m, DPlrMtx *A, DPlrMtx *B);

/* ---- compression --------------------------------------------------------- */

/* Replaces dplrd2gep, reference src/dplrd2p.c:187.  Compresses the dense
 * host m x m matrix A (leading dimension lda) into B with nlvl levels
 * (m divisible by 2^nlvl) and absolute error eps split as the reference
 * splits it (eps/4 per level, eps/2 at the last one): each off-diagonal
 * block keeps its singular triplets with sig >= its share of eps.  The SVDs
 * of one level run as one batched call on the GPU.  *info = 0 on success,
 * -1 on an argument error. */
void dplrd2gep(int m, double *A, int lda, DPlrMtx *B, double eps, int nlvl, int *info);

/* ---- helpers (host) ---------------------------------------------------- */

/* Replaces dplrgep2d, reference src/dplrp2d.c:98.  Dense expansion of the
 * m x m PLR matrix into B (leading dimension ldb >= m). */
void dplrgep2d(int m, DPlrMtx *A, double *B, int ldb);

/* Replaces dplrrnk, reference src/dplrrnk.c:22.  ranks[l] = max rank of the
 * off-diagonal blocks at depth l (caller initialises ranks to -1). */
void dplrrnk(int m, DPlrMtx *A, int *ranks);

/* Replaces dplrnz, reference src/dplrnz.c:17.  Stored entries of the tree. */
long int dplrnz(int m, DPlrMtx *A);

/* ---- hot path (GPU) ---------------------------------------------------- */

/* Replaces dplrgepmv, reference src/dplrpmv.c:262 (workers dplrpmv :113,
 * dplrptmv :188, dplrlrmv :19).
 *   y := alpha*A*x + beta*y   (trans 'N' or 'n')
 *   y := alpha*A^T*x + beta*y (any other character)
 * When *alpha == 0 the reference scales y by beta twice (y := beta*beta*y,
 * or y := 0 when beta == 0); this library does the same.  incx and incy
 * must be positive. */
void dplrgepmv(char trans, int m, double *alpha, DPlrMtx *A,
               double *x, int incx, double *beta, double *y, int incy);

/* Replaces dplrgepdmm, reference src/dplrpdmm.c:454 (workers :177, :253,
 * :320, :387, dplrldmm :23).
 *   C := alpha*op(A)*op(B) + beta*C, A m x m PLR, op(B) m x n, C m x n.
 * op(X) is X for 'N'/'n' and X^T for any other character.  ldc >= m. */
void dplrgepdmm(char transa, char transb, int m, int n, double *alpha,
                DPlrMtx *A, double *B, int ldb, double *beta,
                double *C, int ldc);

/* Replaces dplrgeprk, reference src/dplrprk.c:21.
 *   A := A + alpha*u*diag(s)*v^T   (u, v: m x k; s: k entries, stride incs)
 * Low-rank blocks grow by k normalised columns; leaves get a dense update.
 * Columns whose scale falls below the per-level share of eps get sig = 0. */
void dplrgeprk(int m, int k, double *alpha, double *u, int ldu, double *v,
               int ldv, double *s, int incs, DPlrMtx *A, double eps);

/* Replaces dplrgeppmm, reference src/dplrppmm.c:998 (workers dplrppmm
 * :492, dplrptpmm :617, dplrpptmm :742, dplrptptmm :868).
 *   C := alpha*op(A)*op(B) + beta*C for m x m PLR matrices of one uniform
 *   structure; op(X) is X for 'N'/'n' and X^T otherwise.
 * C keeps the reference's structure: each off-diagonal block holds beta
 * times its columns, then the products with the diagonal subtrees, then one
 * low-rank correction per ancestor (added by dplrgeprk); columns normalised,
 * with sig = 0 below eps/4 per level.  The subtree products of one level
 * run as one product on the GPU.  C must not be A or B. */
void dplrgeppmm(char transa, char transb, int m, double *alpha, DPlrMtx *A, DPlrMtx *B,
                double *beta, DPlrMtx *C, double eps);

/* Replaces dplrgeinv, reference src/dplrinv.c:93.
 *   B := inv(A) as an explicit PLR matrix with the reference's structure:
 *   at depth l the off-diagonal ranks are the Woodbury rank plus one
 *   correction block per ancestor, columns unit-norm, sig >= 0.
 * *info is -1 while working, 0 on success, and > 0 when a leaf or a
 * capacitance matrix is singular.  Argument errors return before *info is
 * touched, as in the reference. */
void dplrgeinv(int m, DPlrMtx *A, DPlrMtx *B, double eps, int *info);

#ifdef __cplusplus
}
#endif

#endif /* DPLR_AMD_DPLR_H */
