// Multi-RHS PLR products on FP64 MFMA (v_mfma_f64_16x16x4_f64).
//
// Lane layout of the instruction on gfx950 (measured): with l the lane,
//   A (16x4):  lane l holds A[l%16][l/16]
//   B (4x16):  lane l holds B[l/16][l%16]
//   D (16x16): lane l holds D[l/16 + 4v][l%16], v = 0..3.
//
// mm_reduce: per (level, segment, row tile, 32-rank group, 64-column group)
//   one wavefront computes T = P|_tile^T X|_tile.  The MFMA K dimension runs
//   over rows: in step t lane group g = l/16 contributes row rb + 16g + t, so
//   each lane streams 16 consecutive doubles of one column per 64-row block.
//   A tile that covers its whole segment writes the finished coefficients;
//   otherwise it writes a partial that mm_finish sums in a fixed order.
// mm_expand: one wavefront per 16 rows x 64 columns computes
//   Y^T = beta Y^T + alpha X_leaf^T D^T + sum_d W_d^T Q_d^T, so the D output
//   rows are RHS columns and its columns are matrix rows: lane l%16 walks 16
//   consecutive matrix rows and the stores are contiguous.
#include <cstring>

#include "plr_common.h"
#include "plr_mm.h"

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_mm_reduce(MmTab a, const double *__restrict__ X, int ldx,
                                                       double alpha, int trans, int raw,
                                                       double *__restrict__ part,
                                                       double *__restrict__ w, size_t wstride)
{
    mm_reduce_body<NW>(a, blockIdx.x, X, ldx, alpha, trans, raw, part, w, wstride);
}

template <bool WT>
__global__ __launch_bounds__(256) void k_mm_reduce_fused(MmTab a, const double *__restrict__ X,
                                                         int ldx, double alpha, int trans, int raw,
                                                         double *__restrict__ part,
                                                         double *__restrict__ w, size_t wstride)
{
    mm_reduce_fused_body<WT>(a, blockIdx.x, blockIdx.y, X, ldx, alpha, trans, raw, part, w, wstride);
}

__global__ __launch_bounds__(64 * MMF2_NW) void k_mm_reduce_fused2(MmTab a, const double *__restrict__ X,
                                                                int ldx, double alpha,
                                                                double *__restrict__ part,
                                                                double *__restrict__ w, size_t wstride,
                                                                MmSplit sp)
{
    extern __shared__ double xs_dyn[];
    int nc = -1;
    const int rg = blockIdx.x, cg = blockIdx.y;
    if (sp.ncg0 > 0) {
        if (cg >= sp.ncg0) {
            X = mm_split_base(sp.X1, sp.ncg0, sp.ldx1);
            ldx = sp.ldx1;
            alpha = sp.alpha1;
            nc = sp.ncg0 * MM_CG + sp.nc1;
        } else {
            nc = sp.nc0;
        }
    }
    mm_reduce_fused2_body(a, rg, cg, X, ldx, alpha, part, w, wstride, xs_dyn, nc);
}

template <bool WT>
__global__ __launch_bounds__(256) void k_mm_finish(MmTab a, double alpha, int trans, int raw,
                                                   const double *__restrict__ part,
                                                   double *__restrict__ w, size_t wstride,
                                                   MmSplit sp)
{
    mm_finish_body<WT>(a, blockIdx.x, alpha, trans, raw, part, w, wstride, sp.ncg0, sp.alpha1);
}

template <int TR>
__global__ __launch_bounds__(256, 2) void k_mm_expand2(MmTab a, const double *__restrict__ D,
                                                    const double *__restrict__ X, int ldx,
                                                    double alpha, double beta,
                                                    const double *__restrict__ w, int wtc,
                                                    const double *C, int ldc, double *Y, int ldy,
                                                    double *__restrict__ bmax);

template <int TR>
__global__ __launch_bounds__(64 * ME3_NW) void k_mm_expand3(MmTab a, const double *__restrict__ D,
                                                        const double *__restrict__ X, int ldx,
                                                        double alpha, double beta,
                                                        const double *__restrict__ w, int wtc,
                                                        const double *C, int ldc, double *Y,
                                                        int ldy, double *__restrict__ bmax,
                                                        MmSplit sp)
{
    extern __shared__ double lds_dyn[];
    int nc = -1;
    const int rg = blockIdx.x, cg = blockIdx.y;
    if (sp.ncg0 > 0) {
        if (cg >= sp.ncg0) {
            X = mm_split_base(sp.X1, sp.ncg0, sp.ldx1);
            ldx = sp.ldx1;
            C = mm_split_base(sp.C1, sp.ncg0, sp.ldc1);
            ldc = sp.ldc1;
            Y = mm_split_base(sp.Y1, sp.ncg0, sp.ldy1);
            ldy = sp.ldy1;
            alpha = sp.alpha1;
            beta = sp.beta1;
            nc = sp.ncg0 * MM_CG + sp.nc1;
        } else {
            nc = sp.nc0;
        }
    }
    mm_expand3_body<TR>(a, rg, cg, D, X, ldx, alpha, beta, w, wtc, C, ldc, Y, ldy, bmax, lds_dyn,
                        nc);
}

// The expansion over the table's levels (and the leaf when D is given):
// the third form when its staged ranks fit in LDS, else the second.
static constexpr size_t ME3_LDS_MAX = 150 * 1024;

template <int TR>
static int mm_expand_launch(dplrh_plan *p, const MmTab &t, const double *D, const double *X,
                            int ldx, double alpha, double beta, const double *w, int wtc,
                            const double *C, int ldc, double *Y, int ldy, double *bmax,
                            hipStream_t st, const MmSplit &sp = MmSplit{})
{
    int rows = 0;
    for (int d = 0; d < t.L && t.seglen[d] >= ME3_RB; ++d) rows += t.R[d];
    const size_t lds = me3_lds(rows) * sizeof(double);
    if (p->N % ME3_RB == 0 && lds <= ME3_LDS_MAX) {
        static bool attr = false;
        if (!attr) {
            PLR_HIP(hipFuncSetAttribute((const void *)k_mm_expand3<TR>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)ME3_LDS_MAX));
            attr = true;
        }
        hipLaunchKernelGGL(k_mm_expand3<TR>, dim3(p->N / ME3_RB, t.ncg), dim3(64 * ME3_NW), lds, st,
                           t, D, X, ldx, alpha, beta, w, wtc, C, ldc, Y, ldy, bmax, sp);
    } else {
        if (sp.ncg0 > 0) {
            plr_set_error("split product needs the third-form expansion");
            return -3;
        }
        hipLaunchKernelGGL(k_mm_expand2<TR>, dim3(p->N / ME_RB, t.ncg), dim3(256), 0, st, t, D, X,
                           ldx, alpha, beta, w, wtc, C, ldc, Y, ldy, bmax);
    }
    PLR_HIP(hipGetLastError());
    return 0;
}

template <int TR>
__global__ __launch_bounds__(256, 2) void k_mm_expand2(MmTab a, const double *__restrict__ D,
                                                    const double *__restrict__ X, int ldx,
                                                    double alpha, double beta,
                                                    const double *__restrict__ w, int wtc,
                                                    const double *C, int ldc, double *Y, int ldy,
                                                    double *__restrict__ bmax)
{
    mm_expand2_body<TR>(a, blockIdx.x, blockIdx.y, D, X, ldx, alpha, beta, w, wtc, C, ldc, Y, ldy,
                        bmax);
}

__global__ __launch_bounds__(256) void k_mm_expand(MmTab a, const double *__restrict__ D,
                                                   const double *__restrict__ X, int ldx,
                                                   double alpha, double beta, int trans,
                                                   const double *__restrict__ w, size_t wstride,
                                                   double *__restrict__ Y, int ldy)
{
    mm_expand_body(a, blockIdx.x, blockIdx.y, D, X, ldx, alpha, beta, trans, w, wstride, Y, ldy);
}

// ---------------------------------------------------------------------------
// host side


// Leaf orders 64 * 2^k: every segment is 64 rows or a multiple of 128 (and
// of 256 from 256 on), which the row groups of the MFMA kernels assume.
int plr_mm_ok(const dplrh_plan *p)
{
    return p->b >= 64 && (p->b & (p->b - 1)) == 0 && p->N >= 128;
}

static int mm_part(dplrh_plan *p, size_t blocks)
{
    const size_t need = blocks * MM_BLK;
    if (need <= p->tmp_cap) return 0;
    ++p->ws_gen;
    if (p->tmp) (void)hipFree(p->tmp);
    p->tmp = nullptr;
    p->tmp_cap = 0;
    PLR_HIP(hipMalloc(&p->tmp, need * sizeof(double)));
    p->tmp_cap = need;
    return 0;
}

int plr_mm_part(dplrh_plan *p, size_t blocks) { return mm_part(p, blocks); }

// C[c * wstride + w_off[d] + s * R_d + k] = sum over segment s of level e of
// V_d[k][i] X[c][i], for every d < e at once: the per-level reduction with a
// table whose entries are the coarser levels' V, each cut into level e's
// segments (rows s * seg_e .. +seg_e), over ncols <= 64 columns of X.
// w_off (e + 1 entries) receives the offsets, *wstride their total.
int plr_mm_cross(dplrh_plan *p, int e, const double *X, int ldx, int ncols, double *C,
                 int *w_off, size_t *wstride, hipStream_t st)
{
    if (e <= 0 || ncols <= 0 || ncols > MM_CG) {
        plr_set_error("plr_mm_cross: level %d, %d columns", e, ncols);
        return -5;
    }
    MmTab t;
    memset(&t, 0, sizeof(t));
    const int N = p->N, seg = N >> (e + 1), nseg = 2 << e;
    const int T = seg < MM_TMAX ? seg : MM_TMAX, tps = seg / T;
    t.L = e;
    t.N = N;
    t.b = p->b;
    t.ncols = ncols;
    t.ncg = 1;
    int task = 0, part = 0, fin = 0, woff = 0;
    for (int d = 0; d < e; ++d) {
        const int R = p->R[d], nkg = (R + MM_KG - 1) / MM_KG;
        t.R[d] = R;
        t.seglen[d] = seg;
        t.T[d] = T;
        t.tps[d] = tps;
        t.nkg[d] = nkg;
        t.task_off[d] = task;
        task += nkg * nseg * tps;
        t.part_off[d] = part;
        if (tps > 1) part += nkg * nseg * tps;
        t.fin_off[d] = fin;
        if (tps > 1) fin += nkg * nseg;
        t.w_off[d] = woff;
        w_off[d] = woff;
        woff += nseg * R;
        t.P[d] = p->V + p->uoff[d];
    }
    t.task_off[e] = task;
    t.part_off[e] = part;
    t.fin_off[e] = fin;
    t.w_off[e] = woff;
    w_off[e] = woff;
    *wstride = (size_t)woff;
    int rc = mm_part(p, (size_t)part);
    if (rc) return rc;
    if (task > 0) {
        if (T >= 256)
            hipLaunchKernelGGL(k_mm_reduce<4>, dim3(task), dim3(256), 0, st, t, X, ldx, 1.0, 0, 1,
                               p->tmp, C, (size_t)woff);
        else
            hipLaunchKernelGGL(k_mm_reduce<1>, dim3(task), dim3(64), 0, st, t, X, ldx, 1.0, 0, 1,
                               p->tmp, C, (size_t)woff);
        PLR_HIP(hipGetLastError());
    }
    if (fin > 0) {
        hipLaunchKernelGGL(k_mm_finish<false>, dim3(fin * MM_FCH), dim3(256), 0, st, t, 1.0, 0, 1,
                           p->tmp, C, (size_t)woff, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    return 0;
}

// As plr_mm_cross through the all-level reduction of the product (one pass
// over X per 256-row group, every table entry from the staged rows), with the
// output rank-major: C[(w_off[d] + s * R_d + k) * 64 + c] for c < ncols <= 64.
// cap: doubles C holds.  Returns 1, launching nothing, when the shape or X
// does not suit that kernel, or C is too small.
// Doubles of plr_mm_cross_rm's output and of its partial sums for level e
// (64 columns); 1 when the shape takes plr_mm_cross instead.
int plr_mm_cross_rm_size(const dplrh_plan *p, int e, size_t *c_doubles, size_t *part_doubles)
{
    const bool leaves_ok = p->b >= 64 && (p->b & (p->b - 1)) == 0;
    if (e <= 0 || p->N % MMF_RB != 0 || !leaves_ok || p->force_valu) return 1;
    const int seg = p->N >> (e + 1), nseg = 2 << e;
    const int T = seg < MM_TMAX ? seg : MM_TMAX, tps = seg / T;
    size_t part = 0, woff = 0;
    for (int d = 0; d < e; ++d) {
        const int R = p->R[d], nkg = (R + MM_KG - 1) / MM_KG;
        if (tps > 1) part += (size_t)nkg * nseg * tps;
        woff += (size_t)nseg * R;
    }
    *c_doubles = woff * MM_CG;
    *part_doubles = part * MM_BLK;
    return 0;
}

// With `pbuf` the partial sums go there (1 returned, nothing launched, when
// they need more than pcap doubles), else to the plan's shared buffer.
int plr_mm_cross_rm(dplrh_plan *p, int e, const double *X, int ldx, int ncols, double *C,
                    size_t cap, int *w_off, hipStream_t st, double *pbuf, size_t pcap)
{
    if (e <= 0 || ncols <= 0 || ncols > MM_CG) {
        plr_set_error("plr_mm_cross_rm: level %d, %d columns", e, ncols);
        return -5;
    }
    const bool leaves_ok = p->b >= 64 && (p->b & (p->b - 1)) == 0;
    if (p->N % MMF_RB != 0 || !leaves_ok || ldx % 2 != 0 || ((uintptr_t)X & 15) != 0 ||
        p->force_valu)
        return 1;
    MmTab t;
    memset(&t, 0, sizeof(t));
    const int N = p->N, seg = N >> (e + 1), nseg = 2 << e;
    const int T = seg < MM_TMAX ? seg : MM_TMAX, tps = seg / T;
    t.L = e;
    t.N = N;
    t.b = p->b;
    t.ncols = ncols;
    t.ncg = 1;
    t.raw = 1;
    int task = 0, part = 0, fin = 0, woff = 0;
    for (int d = 0; d < e; ++d) {
        const int R = p->R[d], nkg = (R + MM_KG - 1) / MM_KG;
        t.R[d] = R;
        t.seglen[d] = seg;
        t.T[d] = T;
        t.tps[d] = tps;
        t.nkg[d] = nkg;
        t.task_off[d] = task;
        task += nkg * nseg * tps;
        t.part_off[d] = part;
        if (tps > 1) part += nkg * nseg * tps;
        t.fin_off[d] = fin;
        if (tps > 1) fin += nkg * nseg;
        t.w_off[d] = woff;
        w_off[d] = woff;
        woff += nseg * R;
        t.P[d] = p->V + p->uoff[d];
    }
    t.task_off[e] = task;
    t.part_off[e] = part;
    t.fin_off[e] = fin;
    t.w_off[e] = woff;
    w_off[e] = woff;
    if ((size_t)woff * MM_CG > cap) return 1;
    if (pbuf) {
        if ((size_t)part * MM_BLK > pcap) return 1;
    } else {
        int rc = mm_part(p, (size_t)part);
        if (rc) return rc;
        pbuf = p->tmp;
    }
    if (task > 0) {
        static bool attr = false;
        if (!attr) {
            PLR_HIP(hipFuncSetAttribute((const void *)k_mm_reduce_fused2,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
            attr = true;
        }
        hipLaunchKernelGGL(k_mm_reduce_fused2, dim3(N / MMF_RB, 1), dim3(64 * MMF2_NW),
                           (size_t)MMF_RB * MMF2_XS * sizeof(double), st, t, X, ldx, 1.0, pbuf,
                           C, (size_t)MM_CG, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    if (fin > 0) {
        hipLaunchKernelGGL(k_mm_finish<true>, dim3(fin * MM_FCH), dim3(256), 0, st, t, 1.0, 0, 1,
                           pbuf, C, (size_t)MM_CG, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    return 0;
}

// Reduce-finish-expand over all levels for ncols columns (the multi-RHS
// product): Y := alpha*op(A)*X + beta*Y.  w must hold soff[L] doubles per
// column (stride wstride).
int plr_mm_apply(dplrh_plan *p, int trans, int ncols, double alpha, double beta,
                 const double *X, int ldx, double *Y, int ldy, double *w, size_t wstride,
                 hipStream_t st)
{
    return plr_mm_apply_c(p, trans, ncols, alpha, beta, X, ldx, nullptr, 0, Y, ldy, w, wstride,
                          nullptr, st);
}

// As plr_mm_apply with the beta term read from C (ldc) instead of Y, and,
// on the fused path, the per-(column, 128-row block) maxima of |Y| in bmax.
// Returns 1, launching nothing, when C or bmax is given and the shape does
// not take the fused path.
int plr_mm_apply_c(dplrh_plan *p, int trans, int ncols, double alpha, double beta,
                   const double *X, int ldx, const double *C, int ldc, double *Y, int ldy,
                   double *w, size_t wstride, double *bmax, hipStream_t st)
{
    const double *P[PLR_MAXL], *Q[PLR_MAXL], *S[PLR_MAXL];
    for (int d = 0; d < p->L; ++d) {
        const double *Ud = p->U + p->uoff[d], *Vd = p->V + p->uoff[d];
        P[d] = trans ? Ud : Vd;
        Q[d] = trans ? Vd : Ud;
        S[d] = p->S + p->soff[d];
    }
    MmTab t;
    mm_tab(p, 0, p->L, P, Q, S, ncols, 0, &t);
    int rc = mm_part(p, (size_t)t.part_off[t.L]);
    if (rc) return rc;
    // the fused path keeps the coefficients rank-major (w holds soff[L] rows
    // of t.ncg * 64 columns: the caller sizes it for whole 64-column groups)
    // and needs leaves of a power of two >= 64 rows (whole 64-row steps in
    // every segment, no leaf across a 128-row expansion block)
    const bool leaves_ok = p->L == 0 || (p->b >= 64 && (p->b & (p->b - 1)) == 0);
    const bool fused = p->N % MMF_RB == 0 && leaves_ok && ldx % 2 == 0 &&
                       ((uintptr_t)X & 15) == 0 && !p->force_valu && w == p->w &&
                       t.ncg * MM_CG <= p->wk_cols;
    if (!fused && (C || bmax)) return 1;
    if (fused) {
        const int wtc = t.ncg * MM_CG;
        if (t.task_off[t.L] > 0) {
            static bool attr2 = false;
            if (!attr2) {
                PLR_HIP(hipFuncSetAttribute((const void *)k_mm_reduce_fused2,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
                attr2 = true;
            }
            hipLaunchKernelGGL(k_mm_reduce_fused2, dim3(p->N / MMF_RB, t.ncg), dim3(64 * MMF2_NW),
                               (size_t)MMF_RB * MMF2_XS * sizeof(double), st, t, X, ldx, alpha, p->tmp,
                               w, (size_t)wtc, MmSplit{});
            PLR_HIP(hipGetLastError());
        }
        if (t.fin_off[t.L] > 0) {
            hipLaunchKernelGGL(k_mm_finish<true>, dim3(t.fin_off[t.L] * MM_FCH), dim3(256), 0, st, t,
                               alpha, trans, 0, p->tmp, w, (size_t)wtc, MmSplit{});
            PLR_HIP(hipGetLastError());
        }
        return trans ? mm_expand_launch<1>(p, t, p->D, X, ldx, alpha, beta, w, wtc, C ? C : Y,
                                           C ? ldc : ldy, Y, ldy, bmax, st)
                     : mm_expand_launch<0>(p, t, p->D, X, ldx, alpha, beta, w, wtc, C ? C : Y,
                                           C ? ldc : ldy, Y, ldy, bmax, st);
    }
    if (t.task_off[t.L] > 0) {
        hipLaunchKernelGGL(k_mm_reduce<1>, dim3(t.task_off[t.L]), dim3(64), 0, st, t, X, ldx, alpha,
                           trans, 0, p->tmp, w, wstride);
        PLR_HIP(hipGetLastError());
    }
    if (t.fin_off[t.L] > 0) {
        hipLaunchKernelGGL(k_mm_finish<false>, dim3(t.fin_off[t.L] * MM_FCH), dim3(256), 0, st, t, alpha,
                           trans, 0, p->tmp, w, wstride, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    dim3 grid((p->N + 63) / 64, t.ncg);
    hipLaunchKernelGGL(k_mm_expand, grid, dim3(256), 0, st, t, p->D, X, ldx, alpha, beta, trans, w,
                       wstride, Y, ldy);
    PLR_HIP(hipGetLastError());
    return 0;
}

// The mixed step's two products in one pass over the factors:
// Y0 := A X0 (nc0 <= 64 columns) and Y1 := C1 - A X1 (nc1 <= 64 columns, the
// refinement's residual), with the maxima of |Y1| per (column, 128-row
// block) in bmax from column 64 on (bmax holds 128 columns).  Returns 1,
// launching nothing, when the shape or the operands do not suit the fused
// kernels (the caller then makes the two products apart).
int plr_mm_apply_split(dplrh_plan *p, int nc0, const double *X0, int ldx0, double *Y0, int ldy0,
                       int nc1, const double *X1, int ldx1, const double *C1, int ldc1, double *Y1,
                       int ldy1, double *bmax, hipStream_t st)
{
    if (nc0 < 1 || nc0 > MM_CG || nc1 < 1 || nc1 > MM_CG) return 1;
    const bool leaves_ok = p->L == 0 || (p->b >= 64 && (p->b & (p->b - 1)) == 0);
    const bool al = ((((uintptr_t)X0 | (uintptr_t)X1) & 15) == 0) && ldx0 % 2 == 0 && ldx1 % 2 == 0;
    if (p->N % MMF_RB != 0 || p->N % ME3_RB != 0 || !leaves_ok || !al || p->force_valu) return 1;
    int rc = plr_workspace(p, 2 * MM_CG);
    if (rc) return rc;
    const double *P[PLR_MAXL], *Q[PLR_MAXL], *S[PLR_MAXL];
    for (int d = 0; d < p->L; ++d) {
        P[d] = p->V + p->uoff[d];
        Q[d] = p->U + p->uoff[d];
        S[d] = p->S + p->soff[d];
    }
    MmTab t;
    mm_tab(p, 0, p->L, P, Q, S, MM_CG + nc1, 0, &t);
    if ((rc = mm_part(p, (size_t)t.part_off[t.L]))) return rc;
    const int wtc = t.ncg * MM_CG;
    MmSplit sp{};
    sp.ncg0 = 1;
    sp.nc0 = nc0;
    sp.nc1 = nc1;
    sp.X1 = X1;
    sp.C1 = C1;
    sp.Y1 = Y1;
    sp.ldx1 = ldx1;
    sp.ldc1 = ldc1;
    sp.ldy1 = ldy1;
    sp.alpha1 = -1.0;
    sp.beta1 = 1.0;
    if (t.task_off[t.L] > 0) {
        static bool attr2 = false;
        if (!attr2) {
            PLR_HIP(hipFuncSetAttribute((const void *)k_mm_reduce_fused2,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
            attr2 = true;
        }
        hipLaunchKernelGGL(k_mm_reduce_fused2, dim3(p->N / MMF_RB, t.ncg), dim3(64 * MMF2_NW),
                           (size_t)MMF_RB * MMF2_XS * sizeof(double), st, t, X0, ldx0, 1.0, p->tmp,
                           p->w, (size_t)wtc, sp);
        PLR_HIP(hipGetLastError());
    }
    if (t.fin_off[t.L] > 0) {
        hipLaunchKernelGGL(k_mm_finish<true>, dim3(t.fin_off[t.L] * MM_FCH), dim3(256), 0, st, t,
                           1.0, 0, 0, p->tmp, p->w, (size_t)wtc, sp);
        PLR_HIP(hipGetLastError());
    }
    return mm_expand_launch<0>(p, t, p->D, X0, ldx0, 1.0, 0.0, p->w, wtc, Y0, ldy0, Y0, ldy0, bmax,
                               st, sp);
}

// Raw reduction T = P_d^T X per segment of level d (T layout: column c at
// T + c*tstride, segment s at s*R).
int plr_mm_reduce1(dplrh_plan *p, int d, const double *P, const double *X, int ldx, int ncols,
                   double *T, size_t tstride, hipStream_t st)
{
    const double *Q = nullptr, *S = nullptr;
    MmTab t;
    mm_tab(p, d, d + 1, &P, &Q, &S, ncols, 1, &t);
    int rc = mm_part(p, (size_t)t.part_off[1]);
    if (rc) return rc;
    if (t.task_off[1] > 0) {
        if (t.T[0] >= 256)
            hipLaunchKernelGGL(k_mm_reduce<4>, dim3(t.task_off[1]), dim3(256), 0, st, t, X, ldx, 1.0,
                               0, 1, p->tmp, T, tstride);
        else
            hipLaunchKernelGGL(k_mm_reduce<1>, dim3(t.task_off[1]), dim3(64), 0, st, t, X, ldx, 1.0,
                               0, 1, p->tmp, T, tstride);
        PLR_HIP(hipGetLastError());
    }
    if (t.fin_off[1] > 0) {
        hipLaunchKernelGGL(k_mm_finish<false>, dim3(t.fin_off[1] * MM_FCH), dim3(256), 0, st, t, 1.0, 0, 1, p->tmp,
                           T, tstride, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    return 0;
}

// Y += Q_d W per segment of level d (no leaf term), or, with d < 0, the leaf
// term alone: Y := alpha*op(D) X + beta*Y with D given.
int plr_mm_expand1(dplrh_plan *p, int d, const double *Q, const double *W, size_t wstride,
                   const double *D, int trans, const double *X, int ldx, double alpha, double beta,
                   double *Y, int ldy, int ncols, hipStream_t st)
{
    const double *P = nullptr, *S = nullptr;
    MmTab t;
    if (d >= 0)
        mm_tab(p, d, d + 1, &P, &Q, &S, ncols, 1, &t);
    else
        mm_tab(p, 0, 0, &P, &Q, &S, ncols, 1, &t);
    dim3 grid((p->N + 63) / 64, t.ncg);
    hipLaunchKernelGGL(k_mm_expand, grid, dim3(256), 0, st, t, D, X, ldx, alpha, beta, trans, W,
                       wstride, Y, ldy);
    PLR_HIP(hipGetLastError());
    return 0;
}

// T0 of levels [0, d1) over X for the reassociated upper solve:
// T[(soff[d] + s*R_d + k) * ws + c] = V_d|_s^T X|_s (raw), through the
// second-form reduction and its finishing sums.  Returns 1, launching
// nothing, when X or the shape does not suit that kernel.
int plr_mm_reduce_raw(dplrh_plan *p, int d1, const double *X, int ldx, int ncols, double *T,
                      size_t ws, hipStream_t st)
{
    const double *P[PLR_MAXL], *Q[PLR_MAXL], *S[PLR_MAXL];
    for (int d = 0; d < d1; ++d) {
        P[d] = p->V + p->uoff[d];
        Q[d] = p->U + p->uoff[d];
        S[d] = p->S + p->soff[d];
    }
    MmTab t;
    mm_tab(p, 0, d1, P, Q, S, ncols, 0, &t);
    t.raw = 1;
    const bool leaves_ok = p->b >= 64 && (p->b & (p->b - 1)) == 0;
    if (p->N % MMF_RB != 0 || !leaves_ok || ldx % 2 != 0 || ((uintptr_t)X & 15) != 0 ||
        (size_t)t.ncg * MM_CG > ws)
        return 1;
    int rc = mm_part(p, (size_t)t.part_off[t.L]);
    if (rc) return rc;
    if (t.task_off[t.L] > 0) {
        static bool attr2 = false;
        if (!attr2) {
            PLR_HIP(hipFuncSetAttribute((const void *)k_mm_reduce_fused2,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
            attr2 = true;
        }
        hipLaunchKernelGGL(k_mm_reduce_fused2, dim3(p->N / MMF_RB, t.ncg), dim3(64 * MMF2_NW),
                           (size_t)MMF_RB * MMF2_XS * sizeof(double), st, t, X, ldx, 1.0, p->tmp, T,
                           ws, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    if (t.fin_off[t.L] > 0) {
        hipLaunchKernelGGL(k_mm_finish<true>, dim3(t.fin_off[t.L] * MM_FCH), dim3(256), 0, st, t, 1.0,
                           0, 1, p->tmp, T, ws, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    return 0;
}

// Y += sum_{d < d1} Q_d|_s W_d[s] with W rank-major (rows soff[d] + s*R_d + k,
// stride ws) and no sig factor: the expansion of the reassociated upper
// solve (Q = Ut).  Y must be 16-byte aligned with an even ldy.
int plr_mm_expand_w(dplrh_plan *p, int d1, const double *Qbase, const double *W, size_t ws,
                    double *Y, int ldy, int ncols, hipStream_t st)
{
    const double *P[PLR_MAXL], *Q[PLR_MAXL], *S[PLR_MAXL];
    for (int d = 0; d < d1; ++d) {
        P[d] = p->V + p->uoff[d];
        Q[d] = Qbase + p->uoff[d];
        S[d] = p->S + p->soff[d];
    }
    MmTab t;
    mm_tab(p, 0, d1, P, Q, S, ncols, 0, &t);
    t.nosig = 1;
    if (p->N % ME_RB != 0 || (size_t)t.ncg * MM_CG > ws) return 1;
    return mm_expand_launch<0>(p, t, nullptr, Y, ldy, 1.0, 1.0, W, (int)ws, Y, ldy, Y, ldy,
                               nullptr, st);
}

// One level through the third-form expansion: Y := beta*Y + Q_d|_s W[s]
// with W rank-major and complete (w[(s*R_d + k) * ws + c], no sig), or with
// d < 0 the leaf term Y := alpha*D X + beta*Y.  Returns 1, launching
// nothing, when the shape or the operands do not suit that kernel.
int plr_mm_expand_lvl(dplrh_plan *p, int d, const double *Q, const double *W, size_t ws,
                      const double *D, const double *X, int ldx, double alpha, double beta,
                      double *Y, int ldy, int ncols, hipStream_t st)
{
    const double *P = nullptr, *S = nullptr;
    MmTab t;
    if (d >= 0)
        mm_tab(p, d, d + 1, &P, &Q, &S, ncols, 1, &t);
    else
        mm_tab(p, 0, 0, &P, &Q, &S, ncols, 1, &t);
    t.nosig = 1;
    const bool leaves_ok = p->b >= 64 && (p->b & (p->b - 1)) == 0;
    if (p->N % ME3_RB != 0 || !leaves_ok || p->force_valu || ldy % 2 != 0 ||
        ((uintptr_t)Y & 15) != 0)
        return 1;
    if (d < 0 && (!D || ldx % 2 != 0 || ((uintptr_t)X & 15) != 0)) return 1;
    if (d >= 0 && ((size_t)t.ncg * MM_CG > ws || ((uintptr_t)W & 15) != 0)) return 1;
    // the kernel reads one element of w even when no level is staged: with
    // no level it gets X, a valid address whose value it does not use
    return mm_expand_launch<0>(p, t, d < 0 ? D : nullptr, X, ldx, alpha, beta, d < 0 ? X : W,
                               (int)ws, Y, ldy, Y, ldy, nullptr, st);
}

// T[c][s*R_d + k] = sum of the raw partial blocks in p->tmp that a caller
// wrote in mm_reduce_body's layout with tiles of `tmax` rows (level d,
// ncols columns).
int plr_mm_finish_raw(dplrh_plan *p, int d, int ncols, int tmax, double *T, size_t tstride,
                      hipStream_t st)
{
    const double *P = nullptr, *Q = nullptr, *S = nullptr;
    MmTab t;
    mm_tab(p, d, d + 1, &P, &Q, &S, ncols, 1, &t, tmax);
    if (t.fin_off[1] > 0) {
        hipLaunchKernelGGL(k_mm_finish<false>, dim3(t.fin_off[1] * MM_FCH), dim3(256), 0, st, t, 1.0,
                           0, 1, p->tmp, T, tstride, MmSplit{});
        PLR_HIP(hipGetLastError());
    }
    return 0;
}

// Partial blocks (of MM_BLK doubles) the levels [d0, d1) need with tiles of
// `tmax` rows and ncols columns.
size_t plr_mm_part_blocks(const dplrh_plan *p, int d0, int d1, int ncols, int tmax)
{
    size_t need = 0;
    for (int d = d0; d < d1; ++d) {
        const double *P = nullptr, *Q = nullptr, *S = nullptr;
        MmTab t;
        mm_tab(p, d, d + 1, &P, &Q, &S, ncols, 1, &t, tmax);
        if ((size_t)t.part_off[1] > need) need = t.part_off[1];
    }
    return need;
}
