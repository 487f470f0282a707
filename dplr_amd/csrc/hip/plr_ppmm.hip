// Level products of dplrgeppmm (the host recursion is csrc/host/dplr_ppmm.c).
//
// At an H-node of depth d the reference multiplies the diagonal subtrees of
// one operand with the low-rank factors of the other operand's off-diagonal
// blocks (src/dplrppmm.c:275, :289 through dplrgepdmm).  In the level-wise
// layout the U factors of level d tile all N rows with the block of node n
// on its two segments, and the diagonal subtrees below depth d form a
// block-diagonal forest over the same rows, so every such product of one
// level is one product of the forest with an N x R_d panel:
//   PU_d = alpha * op(A)_{>d} * U_d(op(B)),   PV_d = alpha * op(B)_{>d}^T * V_d(op(A)),
// where U_d(op(X)) is X's level-d U panel, or its V panel when X is
// transposed (the ne block of X^T is (X.sw)^T, whose U factor is X.sw's V,
// stored on the same upper segment), and likewise for V_d.
#include <cstdlib>
#include <cstring>

#include "plr_common.h"

// Y := alpha * op(A_{>=d0}) X for n device columns (ldx, ldy), A_{>=d0} the
// leaves and levels d0 .. L-1; the general-shape reduce/expand kernels.
static int gemm_from(dplrh_plan *p, int trans, int d0, int n, double alpha, const double *X,
                     int ldx, double *Y, int ldy, hipStream_t st)
{
    if (n <= 0 || p->N == 0) return 0;
    constexpr int chunk = 64;
    int rc = plr_workspace(p, n < chunk ? n : chunk);
    if (rc) return rc;
    LevelTab t;
    plr_level_tab_from(p, trans, d0, &t);
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int nc = n - c0 < chunk ? n - c0 : chunk;
        if ((rc = plr_launch_reduce(t, X + (size_t)c0 * ldx, ldx, alpha, trans, 0, nc, p->part,
                                    p->part_per_col, p->cnt, p->cnt_per_col, p->w, p->w_per_col,
                                    st)))
            return rc;
        if ((rc = plr_launch_expand(t, p->D, X + (size_t)c0 * ldx, ldx, alpha, 0.0, trans, p->w,
                                    p->w_per_col, Y + (size_t)c0 * ldy, ldy, nc, st)))
            return rc;
    }
    return 0;
}

extern "C" int dplrh_ppmm_levels(dplrh_plan *pa, dplrh_plan *pb, int ta, int tb, double alpha,
                                 double **PU, double **PV, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (pa->N != pb->N || pa->L != pb->L || pa->b != pb->b) {
        plr_set_error("dplrgeppmm: A and B differ in structure");
        return -5;
    }
    const int N = pa->N, L = pa->L;
    int rmax = 0;
    for (int d = 0; d < L; ++d) {
        rmax = pa->R[d] > rmax ? pa->R[d] : rmax;
        rmax = pb->R[d] > rmax ? pb->R[d] : rmax;
    }
    if (L == 0 || rmax == 0) return 0;
    double *Y = nullptr;
    PLR_HIP(hipMalloc(&Y, (size_t)N * rmax * sizeof(double)));
    int rc = 0;
    for (int d = 0; d < L && !rc; ++d) {
        // PU_d: op(A)'s forest below d times op(B)'s level-d U panel
        const int rb = pb->R[d], ra = pa->R[d];
        if (rb > 0) {
            const double *X = (tb ? pb->V : pb->U) + pb->uoff[d];
            rc = gemm_from(pa, ta, d + 1, rb, alpha, X, N, Y, N, st);
            if (!rc && hipMemcpyAsync(PU[d], Y, (size_t)N * rb * sizeof(double),
                                      hipMemcpyDeviceToHost, st) != hipSuccess)
                rc = -10;
            if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = -10;
        }
        // PV_d: op(B)'s forest below d, transposed, times op(A)'s level-d V panel
        if (!rc && ra > 0) {
            const double *X = (ta ? pa->U : pa->V) + pa->uoff[d];
            rc = gemm_from(pb, !tb, d + 1, ra, alpha, X, N, Y, N, st);
            if (!rc && hipMemcpyAsync(PV[d], Y, (size_t)N * ra * sizeof(double),
                                      hipMemcpyDeviceToHost, st) != hipSuccess)
                rc = -10;
            if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = -10;
        }
    }
    if (rc == -10) plr_set_error("dplrgeppmm: level product or its download failed");
    (void)hipFree(Y);
    return rc;
}
