// Shared definitions of the HIP layer: the device plan, level tables passed
// to kernels by value, error reporting.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../dplr_hip.h"

#define PLR_MAXL DPLRH_MAXL

// Per-level table handed to kernels by value (fits the 4 KB argument limit).
struct LevelTab
{
    int L, N, b;
    int R[PLR_MAXL];        // stored rank of level d
    int seglen[PLR_MAXL];   // segment length N >> (d+1)
    int T[PLR_MAXL];        // rows per reduction tile at level d
    int tps[PLR_MAXL];      // reduction tiles per segment
    int task_off[PLR_MAXL + 1];
    int nkg[PLR_MAXL];      // rank groups per level
    int part_off[PLR_MAXL]; // offset of level d in the partial-sum buffer
    int cnt_off[PLR_MAXL + 1]; // offset of level d in the arrival counters
    int w_off[PLR_MAXL + 1];   // offset of level d in the coefficient buffer
    const double *P[PLR_MAXL]; // reduction operand (V for A x, U for A^T x)
    const double *Q[PLR_MAXL]; // expansion operand (U for A x, V for A^T x)
    const double *S[PLR_MAXL];
};

// Factored inverse (recursive Woodbury in product form), FP64:
//   inv(A) = F_0 F_1 ... F_{L-1} inv(D_leaf),  F_d = I - Ut_d S_d inv(K_d) Z_d^T
// with Ut_d = (block-diagonal inverse below level d) * U_d, and for node n of
// level d the capacitance  K = [I, M_bot S_bot; M_top S_top, I],
// M_s = V_d|_s^T Ut_d|_s.  Kinv holds inv(K) per node, (2R_d)^2 each.
struct plr_fact
{
    int done;
    int info;
    double *Dinv;                  // leaves' inverses, N/b tiles of b x b
    double *Ut;                    // Ut, layout of U
    double *Kinv;                  // capacitance inverses
    size_t koff[PLR_MAXL + 1];     // element offset of level d in Kinv
    int *dinfo;                    // per-matrix LU status
    double *work;                  // LU scratch (leaves, capacitances)
    // reassociated upper solve: G_{d,e}[s] = V_d|_s^T Ut_e|_s for d < e < up_d0,
    // per segment s of level e, R_d x R_e column-major at goff[d][e] + s*R_d*R_e
    double *G;
    int up_d0;
    size_t goff[PLR_MAXL][PLR_MAXL];
    // G built beside the level loop: its own stream (default priority),
    // reduction output and partial sums, so it shares no buffer with the loop
    double *gC, *gpart; // gC sized once for the shape
    size_t gpart_cap;
    hipStream_t gst;
    hipEvent_t gev, gdone;
    int gd0;
};

// Single-precision copies of the factors (plr_sp.hip): type 1 FP32, type 2
// BF16 factor matrices (uint16_t bits); sig and inv(K) FP32 in both.  Ut,
// Dinv and Kinv exist once fact is set.
struct plr_sp
{
    int type, fact;
    unsigned gen;
    void *U, *V, *Ut, *D, *Dinv;
    float *S, *Kinv;
};

// Workspace of the per-level solve kernels, sized for `cols` columns.
struct plr_solve_ws
{
    int cols;
    size_t t_per_col, part_per_col, cnt_per_col;
    double *T, *W, *part, *tmp;
    int *cnt;
    size_t tmp_cap;
    // reassociated upper solve: T0 and W (rows soff[d] + s*R + k), the
    // ping-pong subtree sums A, all with row stride ucols (64-column groups)
    int ucols;
    double *uT0, *uW, *uA[2];
};

struct dplrh_plan
{
    int N, L, b;
    int R[PLR_MAXL];
    size_t uoff[PLR_MAXL + 1], soff[PLR_MAXL + 1];
    double *U, *V, *S, *D; // device, FP64, layout of dplr_hip.h
    int *rank;             // device copy of plr_flat::rank
    // workspace of the reduce/expand kernels, sized for wk_cols columns
    int wk_cols;
    double *part, *w;
    int *cnt;
    size_t part_per_col, w_per_col, cnt_per_col;
    double *tmp; // op(B) staging for transb
    size_t tmp_cap;
    plr_fact f;
    plr_solve_ws sw;
    plr_sp *sp[2];  // single-precision factors, by type - 1
    struct dplrh_batch *self_batch; // batch of this plan alone (single-precision ops)
    int force_valu; // testing aid: use the general-shape kernels everywhere
    // HIP graphs of repeated calls (see plr_run_cached)
    unsigned ws_gen;  // bumped whenever plan-owned device memory is reallocated
    hipStream_t cap;  // private capture stream (default priority)
    struct
    {
        unsigned long long key[10];
        unsigned gen;
        int seen;
        hipGraphExec_t exec;
        unsigned long long tick;
    } graphs[16];
    unsigned long long tick;
};

// Run op(stream) on `user`, through a cached HIP graph when the same call
// (key) was made before with the same workspace generation: the first call
// runs eagerly (and sizes the workspace), the second is captured on a
// private stream and launched, later calls replay it.  op must neither
// allocate nor synchronise once the workspace is sized.
int plr_run_cached_impl(dplrh_plan *p, const unsigned long long *key, hipStream_t user,
                        int (*op)(void *ctx, hipStream_t), void *ctx);
template <typename F>
int plr_run_cached(dplrh_plan *p, const unsigned long long *key, hipStream_t user, F &&f)
{
    auto thunk = [](void *c, hipStream_t s) -> int { return (*(F *)c)(s); };
    return plr_run_cached_impl(p, key, user, thunk, (void *)&f);
}
void plr_graphs_clear(dplrh_plan *p);
// C := alpha*op(A)*B + beta*C without the graph cache (for use inside other ops).
int plr_gemm_direct(dplrh_plan *p, int trans, int n, double alpha, double beta,
                    const double *dB, int ldb, double *dC, int ldc, hipStream_t st);

void plr_set_error(const char *fmt, ...);

#define PLR_HIP(call)                                                                 \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            plr_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return -10;                                                               \
        }                                                                             \
    } while (0)

// Fill the level table for A x (trans == 0) or A^T x (trans == 1).
void plr_level_tab(const dplrh_plan *p, int trans, LevelTab *t);
// Make sure the workspace holds ncols columns.
int plr_workspace(dplrh_plan *p, int ncols);
// Doubles of partial sums and coefficients and ints of arrival counters the
// reduce/expand kernels need per column.
void plr_ws_per_col(const dplrh_plan *p, size_t *part, size_t *w, size_t *cnt);
// Level table for level d alone, with explicit operands (P reduced, Q expanded).
void plr_level_tab1(const dplrh_plan *p, int d, const double *P, const double *Q,
                    const double *S, LevelTab *t);
// A table with no levels (leaf term only).
void plr_level_tab0(const dplrh_plan *p, LevelTab *t);
// The levels d0 .. L-1 of A x (trans == 0) or A^T x: with the leaves, the
// block-diagonal forest of the subtrees whose roots lie at depth d0.
void plr_level_tab_from(const dplrh_plan *p, int trans, int d0, LevelTab *t);

int plr_launch_reduce(const LevelTab &t, const double *x, int ldx, double alpha, int trans,
                      int raw, int ncols, double *part, size_t pstride, int *cnt, size_t cstride,
                      double *w, size_t wstride, hipStream_t st);
int plr_launch_expand(const LevelTab &t, const double *D, const double *x, int ldx, double alpha,
                      double beta, int trans, const double *w, size_t wstride, double *y, int ldy,
                      int ncols, hipStream_t st);
void plr_free_fact(dplrh_plan *p);

// Batched FP64 LU with partial pivoting and inverse of `count` n x n
// column-major matrices (K is overwritten by the LU factors); *info = the
// first nonzero dgetrf-style status in matrix order.
int dplrh_lp_lu_inverse(double *K, int n, int count, double *Kinv, int *info, hipStream_t st);
// Single-level building blocks (plr_solve.hip), MFMA when possible.
int plr_lvl_reduce(dplrh_plan *p, int d, const double *P, const double *X, int ldx, int ncols,
                   double *T, size_t tstride, hipStream_t st);
int plr_lvl_expand(dplrh_plan *p, int d, const double *Q, const double *W, size_t wstride,
                   const double *D, int trans, const double *X, int ldx, double alpha,
                   double beta, double *Y, int ldy, int ncols, hipStream_t st);
int plr_lvl_cap(dplrh_plan *p, int d, const double *T, size_t tstride, double *W, int ncols,
                int trans, hipStream_t st);
int plr_solve_reserve(dplrh_plan *p, int cols);

// Mixed-precision factors of a plan (plr_lp.hip).
void plr_lp_release(const dplrh_plan *p);
// Single-precision factors of a plan (plr_sp.hip).
void plr_sp_release(dplrh_plan *p);

// FP64 MFMA kernels (plr_mm.hip); usable when plr_mm_ok(p).
int plr_mm_ok(const dplrh_plan *p);
int plr_mm_apply(dplrh_plan *p, int trans, int ncols, double alpha, double beta,
                 const double *X, int ldx, double *Y, int ldy, double *w, size_t wstride,
                 hipStream_t st);
int plr_mm_apply_c(dplrh_plan *p, int trans, int ncols, double alpha, double beta,
                   const double *X, int ldx, const double *C, int ldc, double *Y, int ldy,
                   double *w, size_t wstride, double *bmax, hipStream_t st);
int plr_mm_reduce1(dplrh_plan *p, int d, const double *P, const double *X, int ldx, int ncols,
                   double *T, size_t tstride, hipStream_t st);
int plr_mm_part(dplrh_plan *p, size_t blocks);
int plr_mm_finish_raw(dplrh_plan *p, int d, int ncols, int tmax, double *T, size_t tstride,
                      hipStream_t st);
size_t plr_mm_part_blocks(const dplrh_plan *p, int d0, int d1, int ncols, int tmax);
int plr_mm_reduce_raw(dplrh_plan *p, int d1, const double *X, int ldx, int ncols, double *T,
                      size_t ws, hipStream_t st);
int plr_mm_cross(dplrh_plan *p, int e, const double *X, int ldx, int ncols, double *C,
                 int *w_off, size_t *wstride, hipStream_t st);
int plr_mm_expand_w(dplrh_plan *p, int d1, const double *Qbase, const double *W, size_t ws,
                    double *Y, int ldy, int ncols, hipStream_t st);
// Reassociated upper levels of the FP64 solve (plr_up.hip).
int plr_up_factor(dplrh_plan *p, hipStream_t st);
int plr_up_begin(dplrh_plan *p);
int plr_up_level(dplrh_plan *p, int e, hipStream_t st);
int plr_up_end(dplrh_plan *p, hipStream_t st);
void plr_up_release(dplrh_plan *p);
int plr_up_solve(dplrh_plan *p, int d0, double *X, int ldx, int ncols, hipStream_t st);
int plr_up_ready(const dplrh_plan *p, int d0);
int plr_up_reserve(dplrh_plan *p, int d0, int ncols);
int plr_up_steps(dplrh_plan *p, int d0, int ncols, hipStream_t st);
int plr_tail_d0(const dplrh_plan *p);
int plr_mm_cross_rm(dplrh_plan *p, int e, const double *X, int ldx, int ncols, double *C,
                    size_t cap, int *w_off, hipStream_t st, double *part = nullptr,
                    size_t part_cap = 0);
int plr_mm_cross_rm_size(const dplrh_plan *p, int e, size_t *c_doubles, size_t *part_doubles);
int plr_mm_apply_split(dplrh_plan *p, int nc0, const double *X0, int ldx0, double *Y0, int ldy0,
                       int nc1, const double *X1, int ldx1, const double *C1, int ldc1, double *Y1,
                       int ldy1, double *bmax, hipStream_t st);
int plr_mm_expand_lvl(dplrh_plan *p, int d, const double *Q, const double *W, size_t ws,
                      const double *D, const double *X, int ldx, double alpha, double beta,
                      double *Y, int ldy, int ncols, hipStream_t st);
int plr_mm_expand1(dplrh_plan *p, int d, const double *Q, const double *W, size_t wstride,
                   const double *D, int trans, const double *X, int ldx, double alpha, double beta,
                   double *Y, int ldy, int ncols, hipStream_t st);
