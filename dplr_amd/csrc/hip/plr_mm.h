// FP64 MFMA building blocks of the multi-RHS products (device bodies shared
// by the single-plan kernels and the batched kernels), and their tables.
#pragma once
#include <cstdint>

#include <cstring>

#include "plr_common.h"

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static constexpr int MM_KG = 32;   // ranks per task (2 MFMA tiles)
static constexpr int MM_CG = 64;   // columns per task (4 MFMA tiles)
static constexpr int MM_TMAX = 256; // rows per reduction task
static constexpr int MM_BLK = MM_KG * MM_CG;

struct MmTab
{
    int L, N, b, ncols, ncg;
    int R[PLR_MAXL], seglen[PLR_MAXL], T[PLR_MAXL], tps[PLR_MAXL], nkg[PLR_MAXL];
    int task_off[PLR_MAXL + 1];
    int part_off[PLR_MAXL + 1]; // in blocks of MM_BLK doubles
    int w_off[PLR_MAXL + 1];
    int fin_off[PLR_MAXL + 1];  // finishing tasks (segments x groups of levels with tps > 1)
    int sx; // the reduction stores alpha * t; the expansion multiplies by sig when staging
    int raw;   // second-form reduction: T of the segment itself, unscaled
    int nosig; // second-form expansion: no sig factor (W is complete)
    const double *P[PLR_MAXL], *Q[PLR_MAXL], *S[PLR_MAXL];
};

// Two operand sets in one multi-RHS product (the mixed step's apply and its
// first residual share the factor reads): column groups [0, ncg0) take the
// kernel's own X, C, Y, alpha, beta and nc0 columns; the groups from ncg0 on
// take the second set, its column c at X1 + (c - 64 ncg0) * ldx1, and nc1
// columns.  ncg0 = 0: one operand set.
struct MmSplit
{
    int ncg0, nc0, nc1;
    const double *X1, *C1;
    double *Y1;
    int ldx1, ldc1, ldy1;
    double alpha1, beta1;
};

// the base that puts column 64 * cg0 of a split product's second part at p;
// formed by integer arithmetic, which hides that it points to global
// memory, so it is cast through a global pointer (else every load through
// X, C or Y in the kernels that select it would become a flat load)
template <typename T>
__device__ __forceinline__ T *mm_split_base(T *p, int cg0, int ld)
{
    typedef __attribute__((address_space(1))) T GT;
    return (T *)(GT *)((uintptr_t)p - (uintptr_t)cg0 * MM_CG * (uintptr_t)ld * sizeof(double));
}

// ---------------------------------------------------------------------------

// NW waves per work-group split the rows of a tile (alternate 64-row blocks)
// and are combined in LDS; NW = 1 when there are tasks enough to fill the
// GPU (every level at once), NW = 4 for one level of long segments.
template <int NW>
__device__ __forceinline__ void mm_reduce_body(const MmTab &a, int task, const double *__restrict__ X,
                                               int ldx, double alpha, int trans, int raw,
                                               double *__restrict__ part, double *__restrict__ w,
                                               size_t wstride)
{
    __shared__ double red[NW > 1 ? 2 : 1][NW > 1 ? MM_BLK : 1];
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4, wv = threadIdx.x >> 6;
    int d = 0;
    while (task >= a.task_off[d + 1]) ++d;
    const int loc = task - a.task_off[d];
    const int nkg = a.nkg[d], ncg = a.ncg, tps = a.tps[d];
    const int kg = loc % nkg;
    const int cg = (loc / nkg) % ncg;
    const int st = loc / (nkg * ncg); // segment * tps + tile
    const int s = st / tps, tile = st - s * tps;
    const int R = a.R[d], N = a.N, seg = a.seglen[d], T = a.T[d];
    const int k0 = kg * MM_KG, c0 = cg * MM_CG;
    const int rbeg = s * seg + tile * T;
    const int rend = min(rbeg + T, (s + 1) * seg);
    const int nct = min(4, (a.ncols - c0 + 15) >> 4); // live 16-column tiles (uniform)

    // per-lane column pointers (zero operand for columns past the end)
    const double *pa[2];
    const double *pb[4];
    bool va[2], vb[4];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
        const int k = k0 + kt * 16 + m;
        va[kt] = k < R;
        pa[kt] = a.P[d] + (size_t)(va[kt] ? k : 0) * N;
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int c = c0 + ct * 16 + m;
        vb[ct] = c < a.ncols;
        pb[ct] = X + (size_t)(vb[ct] ? c : 0) * ldx;
    }
    d4 acc[2][4];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[kt][ct] = d4{0.0, 0.0, 0.0, 0.0};

    // Loads are unconditional (columns past the end read column 0) and are
    // zeroed after they land, so a whole 8-row chunk is in flight at once.
    // The four waves take alternate 64-row blocks of the tile.
    for (int rb = rbeg + 64 * wv; rb < rend; rb += 64 * NW) {
        const int r = rb + 16 * g;
#pragma unroll
        for (int t0 = 0; t0 < 16; t0 += 8) {
            double2 ua[2][4], ub[4][4];
#pragma unroll
            for (int kt = 0; kt < 2; ++kt)
#pragma unroll
                for (int h = 0; h < 4; ++h) ua[kt][h] = *(const double2 *)(pa[kt] + r + t0 + 2 * h);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                if (ct >= nct) continue; // uniform
#pragma unroll
                for (int h = 0; h < 4; ++h) ub[ct][h] = *(const double2 *)(pb[ct] + r + t0 + 2 * h);
            }
#pragma unroll
            for (int tt = 0; tt < 8; ++tt) {
                double av[2], bv[4];
#pragma unroll
                for (int kt = 0; kt < 2; ++kt) {
                    const double v = (tt & 1) ? ua[kt][tt >> 1].y : ua[kt][tt >> 1].x;
                    av[kt] = va[kt] ? v : 0.0;
                }
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    if (ct >= nct) break;
                    const double v = (tt & 1) ? ub[ct][tt >> 1].y : ub[ct][tt >> 1].x;
                    bv[ct] = vb[ct] ? v : 0.0;
                }
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    if (ct >= nct) break;
#pragma unroll
                    for (int kt = 0; kt < 2; ++kt) acc[kt][ct] = mfma64(av[kt], bv[ct], acc[kt][ct]);
                }
            }
        }
    }

    // combine the four waves in a fixed order: (w0 + w2) + (w1 + w3)
    if (NW > 1) {
        const int slot = (wv & 1), base = lane;
        if (wv >= 2) {
#pragma unroll
            for (int kt = 0; kt < 2; ++kt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                    for (int v = 0; v < 4; ++v) red[slot][((kt * 4 + ct) * 4 + v) * 64 + base] = acc[kt][ct][v];
        }
        __syncthreads();
        if (wv < 2) {
#pragma unroll
            for (int kt = 0; kt < 2; ++kt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[kt][ct][v] += red[slot][((kt * 4 + ct) * 4 + v) * 64 + base];
        }
        __syncthreads();
        if (wv == 1) {
#pragma unroll
            for (int kt = 0; kt < 2; ++kt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                    for (int v = 0; v < 4; ++v) red[0][((kt * 4 + ct) * 4 + v) * 64 + base] = acc[kt][ct][v];
        }
        __syncthreads();
        if (wv != 0) return;
#pragma unroll
        for (int kt = 0; kt < 2; ++kt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[kt][ct][v] += red[0][((kt * 4 + ct) * 4 + v) * 64 + base];
    }

    // lane holds T[k0 + 16kt + g + 4v][c0 + 16ct + m]
    if (tps == 1) {
        const int sp = raw ? s : (s ^ 1);
#pragma unroll
        for (int kt = 0; kt < 2; ++kt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int c = c0 + ct * 16 + m;
                if (c >= a.ncols) continue;
                double *wc = w + (size_t)c * wstride + a.w_off[d] + (size_t)sp * R;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = k0 + kt * 16 + g + 4 * v;
                    if (k < R)
                        wc[k] = raw ? acc[kt][ct][v] : alpha * acc[kt][ct][v];
                }
            }
        return;
    }
    // partial block: [c_local][k_local], c_local < 64, k_local < 32
    double *pp = part + ((size_t)a.part_off[d] + ((size_t)st * nkg + kg) * ncg + cg) * MM_BLK;
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                pp[(ct * 16 + m) * MM_KG + kt * 16 + g + 4 * v] = acc[kt][ct][v];
}


// All-level reduction for the multi-RHS product: a work-group of 4 waves owns
// 256 rows x 64 columns of X and every 16-rank tile of every level.  Each
// 64-row block of X is staged once in LDS (its loads in flight during the
// previous block's MFMAs) and serves all levels; wave w takes tiles w, w+4,
// w+8 of a pass (one wave per SIMD, accumulators of three tiles),
// streaming its P rows in half-blocks of 8 rows per lane with the next
// half-block's loads issued before the current one is used.  Levels whose segments end inside the
// work-group's rows write their coefficients when a segment ends; longer
// segments write the partial blocks of mm_reduce_body (tiles of 256 rows),
// summed by mm_finish_body.
static constexpr int MMF_RB = 256, MMF_NW = 4, MMF_TPW = 3;

// WT: coefficients rank-major, w[(w_off[d] + s*R + k) * wstride + c] (wstride
// = the padded column count), and partial blocks [k_local][c_local].
template <bool WT = false>
__device__ __forceinline__ void mm_reduce_fused_body(const MmTab &a, int rbk, int cg,
                                                     const double *__restrict__ X, int ldx,
                                                     double alpha, int trans, int raw,
                                                     double *__restrict__ part,
                                                     double *__restrict__ w, size_t wstride)
{
    __shared__ double xs[64 * 65];
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int wv = threadIdx.x >> 6;
    const int rb0 = rbk * MMF_RB, c0 = cg * MM_CG, N = a.N;
    int ntile = 0;
    for (int d = 0; d < a.L; ++d) ntile += (a.R[d] + 15) / 16;
    // X staging: each thread loads 8 pairs of rows (16 bytes each)
    double2 xr[8];
    auto xload = [&](int r0) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = u * 256 + (int)threadIdx.x, c = e >> 5, pr = e & 31;
            const int cc = min(c0 + c, a.ncols - 1);
            xr[u] = *(const double2 *)(X + (size_t)cc * ldx + r0 + 2 * pr);
        }
    };
    auto xstore = [&](double *sh) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = u * 256 + (int)threadIdx.x, c = e >> 5, pr = e & 31;
            const double on = c0 + c < a.ncols ? 1.0 : 0.0;
            sh[c * 65 + 2 * pr] = xr[u].x * on;
            sh[c * 65 + 2 * pr + 1] = xr[u].y * on;
        }
    };
    for (int pass = 0; pass * MMF_NW * MMF_TPW < ntile; ++pass) {
        int td[MMF_TPW], tk[MMF_TPW];
        bool tv[MMF_TPW];
        const double *pa[MMF_TPW];
        double pon[MMF_TPW];
#pragma unroll
        for (int j = 0; j < MMF_TPW; ++j) {
            const int t = pass * MMF_NW * MMF_TPW + wv + MMF_NW * j;
            tv[j] = t < ntile;
            int d = 0, cum = 0;
            const int tt = tv[j] ? t : 0;
            while (tt >= cum + (a.R[d] + 15) / 16) cum += (a.R[d++] + 15) / 16;
            td[j] = d;
            tk[j] = (tt - cum) * 16;
            const int k = tk[j] + m;
            pon[j] = (tv[j] && k < a.R[d]) ? 1.0 : 0.0;
            pa[j] = a.P[d] + (size_t)min(k, a.R[d] - 1) * N;
        }
        d4 acc[MMF_TPW][4];
#pragma unroll
        for (int j = 0; j < MMF_TPW; ++j)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[j][ct] = d4{0.0, 0.0, 0.0, 0.0};
        double2 pA[MMF_TPW][4], pB[MMF_TPW][4];
        // half-block k: rows rb0 + 64*(k/2) + 16g + 8*(k%2) + 0..7 of every tile
        auto pload = [&](double2 (*pv)[4], int k) {
            const int r = rb0 + 64 * (k >> 1) + 16 * g + 8 * (k & 1);
#pragma unroll
            for (int j = 0; j < MMF_TPW; ++j)
#pragma unroll
                for (int h = 0; h < 4; ++h) pv[j][h] = *(const double2 *)(pa[j] + r + 2 * h);
        };
        auto compute = [&](const double *sh, double2 (*pv)[4], int half) {
#pragma unroll
            for (int tt = 0; tt < 8; ++tt) {
                const int t = 8 * half + tt;
                double bq[4];
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) bq[ct] = sh[(16 * ct + m) * 65 + 16 * g + t];
#pragma unroll
                for (int j = 0; j < MMF_TPW; ++j) {
                    const double av = ((tt & 1) ? pv[j][tt >> 1].y : pv[j][tt >> 1].x) * pon[j];
#pragma unroll
                    for (int ct = 0; ct < 4; ++ct) acc[j][ct] = mfma64(av, bq[ct], acc[j][ct]);
                }
            }
        };
        // tile j's output for the segment ending at row rend of its level
        auto flush = [&](int j, int rend) {
            const int d = td[j], R = a.R[d], seg = a.seglen[d];
            const int s = (rend - 1) / seg;
            const int tps = a.tps[d], T = a.T[d];
            if (tps == 1) {
                // stores only: a load here would drain the prefetch in flight
                const int sp = raw ? s : (s ^ 1);
                const double f = raw ? 1.0 : alpha;
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    const int c = c0 + ct * 16 + m;
                    if (c >= a.ncols) continue;
                    if (WT) {
                        double *wc = w + ((size_t)a.w_off[d] + (size_t)sp * R) * wstride + c;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int k = tk[j] + g + 4 * v;
                            if (k < R) wc[(size_t)k * wstride] = f * acc[j][ct][v];
                        }
                    } else {
                        double *wc = w + (size_t)c * wstride + a.w_off[d] + (size_t)sp * R;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int k = tk[j] + g + 4 * v;
                            if (k < R) wc[k] = f * acc[j][ct][v];
                        }
                    }
                }
            } else {
                const int tile = (rend - T - s * seg) / T;
                const int st = s * tps + tile, kg = tk[j] / MM_KG;
                double *pp = part + ((size_t)a.part_off[d] + ((size_t)st * a.nkg[d] + kg) * a.ncg + cg) * MM_BLK;
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int kl = (tk[j] % MM_KG) + g + 4 * v, cl = ct * 16 + m;
                        pp[WT ? kl * MM_CG + cl : cl * MM_KG + kl] = acc[j][ct][v];
                    }
            }
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[j][ct] = d4{0.0, 0.0, 0.0, 0.0};
        };
        auto segment_ends = [&](int rend) {
#pragma unroll
            for (int j = 0; j < MMF_TPW; ++j)
                if (tv[j] && (rend - rb0) % min(a.seglen[td[j]], MMF_RB) == 0) flush(j, rend);
        };
        __syncthreads(); // the previous pass is done with the buffer
        xload(rb0);
        pload(pA, 0);
        xstore(xs);
        __syncthreads();
        // blocks 0..2: the next block of X and the next half-blocks of P are
        // in flight while the current ones are used (the roles of the two
        // P buffers are the same in every iteration, so nothing is copied)
#pragma unroll 1
        for (int blk = 0; blk < 3; ++blk) {
            xload(rb0 + 64 * (blk + 1));
            pload(pB, 2 * blk + 1);
            compute(xs, pA, 0);
            pload(pA, 2 * blk + 2);
            compute(xs, pB, 1);
            segment_ends(rb0 + 64 * (blk + 1));
            __syncthreads();
            xstore(xs);
            __syncthreads();
        }
        pload(pB, 7);
        compute(xs, pA, 0);
        compute(xs, pB, 1);
        segment_ends(rb0 + 256);
    }
}

// All-level reduction, second form (rank-major output only; with a.raw the
// coefficients of segment s are its own unscaled T, else alpha T goes to
// the partner segment s^1 as the expansion needs it): a work-group
// of 8 waves owns 256 rows x 64 columns; X of its rows is staged in LDS once
// (one barrier), then every wave streams the P rows of its 16-rank tiles
// (tiles w, w+8, w+16 of all levels: 20 at rank 32 x 10 levels, more in
// further passes) with the next quarter-block's loads in flight, and never
// waits on another wave.  A wave runs only its valid tiles (the pass is
// specialised on their number), so waves w and w+4, which share a SIMD,
// together carry 5 of 20 tiles.  Ranks past R in a tile read a duplicate of
// row R-1: they only feed output rows k >= R, which are never stored.
// Every level's segment length must be 64 or 128 or a multiple of 256, and
// the task tile 256 rows where it is longer.
static constexpr int MMF2_NW = 8, MMF2_TPW = 3, MMF2_XS = 72;

template <int NT>
__device__ __forceinline__ void mm_fused2_pass(const MmTab &a, int rb0, int cg, int t0, int wv,
                                               double alpha, double *__restrict__ part,
                                               double *__restrict__ w, size_t wstride,
                                               const double *xs, int NC)
{
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int c0 = cg * MM_CG, N = a.N;
    int td[NT], tk[NT];
    const double *pa[NT];
    d4 acc[NT][4];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int t = t0 + wv + MMF2_NW * j;
        int d = 0, cum = 0;
        while (t >= cum + (a.R[d] + 15) / 16) cum += (a.R[d++] + 15) / 16;
        td[j] = d;
        tk[j] = (t - cum) * 16;
        pa[j] = a.P[d] + (size_t)min(tk[j] + m, a.R[d] - 1) * N;
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[j][ct] = d4{0.0, 0.0, 0.0, 0.0};
    }
    double2 pA[NT][2], pB[NT][2];
    // quarter-block k: rows rb0 + 16k + 8h + 2g + {0,1} (h = 0, 1): the four
    // lane groups of one load read 64 contiguous bytes of a rank row, the
    // two loads its whole 128-byte line
    auto pload = [&](double2 (*pv)[2], int k) {
        const int r = rb0 + 16 * k + 2 * g;
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int h = 0; h < 2; ++h) pv[j][h] = *(const double2 *)(pa[j] + r + 8 * h);
    };
    auto compute = [&](double2 (*pv)[2], int k) {
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            // step tt: k-index g is row 16k + 8(tt/2) + 2g + tt%2
            const int rl = 16 * k + 8 * (tt >> 1) + 2 * g + (tt & 1);
            double bq[4];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) bq[ct] = xs[rl * MMF2_XS + 16 * ct + m];
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const double av = (tt & 1) ? pv[j][tt >> 1].y : pv[j][tt >> 1].x;
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) acc[j][ct] = mfma64(av, bq[ct], acc[j][ct]);
            }
        }
    };
    // tile j's output for the segment ending at row rend of its level
    auto flush = [&](int j, int rend) {
        const int d = td[j], R = a.R[d], seg = a.seglen[d];
        const int s = (rend - 1) / seg;
        const int tps = a.tps[d], T = a.T[d];
        if (tps == 1) {
            // stores only: a load here would drain the prefetch in flight
            const int sp = a.raw ? s : s ^ 1;
            const double f = a.raw ? 1.0 : alpha;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int c = c0 + ct * 16 + m;
                if (c >= NC) continue;
                double *wc = w + ((size_t)a.w_off[d] + (size_t)sp * R) * wstride + c;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = tk[j] + g + 4 * v;
                    if (k < R) wc[(size_t)k * wstride] = f * acc[j][ct][v];
                }
            }
        } else {
            const int tile = (rend - T - s * seg) / T;
            const int st = s * tps + tile, kg = tk[j] / MM_KG;
            double *pp = part + ((size_t)a.part_off[d] + ((size_t)st * a.nkg[d] + kg) * a.ncg + cg) * MM_BLK;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    pp[((tk[j] % MM_KG) + g + 4 * v) * MM_CG + ct * 16 + m] = acc[j][ct][v];
        }
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[j][ct] = d4{0.0, 0.0, 0.0, 0.0};
    };
    auto segment_ends = [&](int rend) {
#pragma unroll
        for (int j = 0; j < NT; ++j)
            if ((rend - rb0) % min(a.seglen[td[j]], MMF_RB) == 0) flush(j, rend);
    };
    // quarter-blocks 0..15: the next one's loads in flight while the
    // current one is used (the two buffers swap roles, nothing is copied;
    // no load is conditional, so no wait covers more than the one buffer)
    pload(pA, 0);
#pragma unroll 1
    for (int k = 0; k < 12; k += 4) {
        pload(pB, k + 1);
        compute(pA, k);
        pload(pA, k + 2);
        compute(pB, k + 1);
        pload(pB, k + 3);
        compute(pA, k + 2);
        pload(pA, k + 4);
        compute(pB, k + 3);
        segment_ends(rb0 + 16 * (k + 4));
    }
    pload(pB, 13);
    compute(pA, 12);
    pload(pA, 14);
    compute(pB, 13);
    pload(pB, 15);
    compute(pA, 14);
    compute(pB, 15);
    segment_ends(rb0 + MMF_RB);
}

__device__ __forceinline__ void mm_reduce_fused2_body(const MmTab &a, int rbk, int cg,
                                                      const double *__restrict__ X, int ldx,
                                                      double alpha, double *__restrict__ part,
                                                      double *__restrict__ w, size_t wstride,
                                                      double *xs, int NC = -1)
{
    // wave-uniform (scalar) so that the per-tile table reads are scalar
    // loads: a vector load there would wait for the P loads in flight
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rb0 = rbk * MMF_RB, c0 = cg * MM_CG;
    if (NC < 0) NC = a.ncols; // columns (a split product gives each part its own)
    int ntile = 0;
    for (int d = 0; d < a.L; ++d) ntile += (a.R[d] + 15) / 16;
    // X of all 256 rows, row-major in LDS (xs[row * 72 + col]: the B reads
    // of one half-wave, 16 columns x 2 rows 2 apart, hit 64 distinct banks),
    // columns c0.. (zero past ncols), then one barrier.  Element pair e:
    // rows 2pr, 2pr+1 of column c, pr = 4*(e>>8) + (e&3), c = 16*((e>>6)&3)
    // + ((e>>2)&15): four lanes read 64 contiguous bytes of one column.
    for (int e0 = 0; e0 < 64 * (MMF_RB / 2); e0 += 8 * 64 * MMF2_NW) {
        double2 xr[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * 64 * MMF2_NW + (int)threadIdx.x;
            const int pr = 4 * (e >> 8) + (e & 3), c = 16 * ((e >> 6) & 3) + ((e >> 2) & 15);
            xr[u] = *(const double2 *)(X + (size_t)min(c0 + c, NC - 1) * ldx + rb0 + 2 * pr);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * 64 * MMF2_NW + (int)threadIdx.x;
            const int pr = 4 * (e >> 8) + (e & 3), c = 16 * ((e >> 6) & 3) + ((e >> 2) & 15);
            const double on = c0 + c < NC ? 1.0 : 0.0;
            xs[(2 * pr) * MMF2_XS + c] = xr[u].x * on;
            xs[(2 * pr + 1) * MMF2_XS + c] = xr[u].y * on;
        }
    }
    __syncthreads();
    for (int t0 = 0; t0 < ntile; t0 += MMF2_NW * MMF2_TPW) {
        const int rest = ntile - t0 - wv; // tiles t0 + wv + 8j exist for 8j < rest
        const int nt = rest <= 0 ? 0 : min(MMF2_TPW, (rest + MMF2_NW - 1) / MMF2_NW);
        if (nt == 3)
            mm_fused2_pass<3>(a, rb0, cg, t0, wv, alpha, part, w, wstride, xs, NC);
        else if (nt == 2)
            mm_fused2_pass<2>(a, rb0, cg, t0, wv, alpha, part, w, wstride, xs, NC);
        else if (nt == 1)
            mm_fused2_pass<1>(a, rb0, cg, t0, wv, alpha, part, w, wstride, xs, NC);
    }
}

// Sum the partial blocks of every segment (fixed order) into the
// coefficients.  One work-group per (level, segment, rank group, column
// group, 64-element chunk of the block): its four waves sum every fourth
// tile with eight loads in flight, then combine through LDS.
static constexpr int MM_FCH = MM_BLK / 64; // chunks per block

template <bool WT = false>
__device__ __forceinline__ void mm_finish_body(const MmTab &a, int bx, double alpha, int trans,
                                               int raw, const double *__restrict__ part,
                                               double *__restrict__ w, size_t wstride,
                                               int ncg0 = 0, double alpha1 = 0.0)
{
    __shared__ double red[4][64];
    const int task = bx / MM_FCH, chunk = bx % MM_FCH;
    int d = 0;
    while (task >= a.fin_off[d + 1]) ++d;
    const int loc = task - a.fin_off[d];
    const int nkg = a.nkg[d], ncg = a.ncg, tps = a.tps[d], R = a.R[d];
    const int kg = loc % nkg, cg = (loc / nkg) % ncg, s = loc / (nkg * ncg);
    if (ncg0 > 0 && cg >= ncg0) alpha = alpha1; // the second part of a split product
    const int el = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int e = chunk * 64 + el;
    const double *pp = part + ((size_t)a.part_off[d] + ((size_t)s * tps * nkg + kg) * ncg + cg) * MM_BLK + e;
    const size_t step = (size_t)nkg * ncg * MM_BLK;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j0 = ph; j0 < tps; j0 += 32) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + 4 * u;
            const double v = pp[(size_t)min(j, tps - 1) * step];
            acc[u] += j < tps ? v : 0.0;
        }
    }
    red[ph][el] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (ph != 0) return;
    const double t = (red[0][el] + red[1][el]) + (red[2][el] + red[3][el]);
    const int kl = WT ? e / MM_CG : e % MM_KG, cl = WT ? e % MM_CG : e / MM_KG;
    const int k = kg * MM_KG + kl, c = cg * MM_CG + cl;
    if (k >= R || c >= a.ncols) return;
    const int sp = raw ? s : (s ^ 1);
    if (WT)
        w[((size_t)a.w_off[d] + (size_t)sp * R + k) * wstride + c] = raw ? t : alpha * t;
    else
        w[(size_t)c * wstride + a.w_off[d] + (size_t)sp * R + k] = raw ? t : alpha * t;
}


// Y^T (64 columns x 16 rows per wavefront) += W^T Q^T over all levels, plus
// the leaf term; four wavefronts per work-group cover 64 rows, which lie in
// one segment of every level.  The A operands (W^T per level, X^T of the
// leaf) are staged in LDS per 64-wide chunk with coalesced loads; the B
// operands (Q^T, D^T) are read from HBM, 16 consecutive rows per lane group.
__device__ __forceinline__ void mm_stage(double *sh, const double *src, size_t cstride, int nc,
                                         int kn, const double *scale = nullptr)
{
    // sh[c*65 + kk] = src[c*cstride + kk] for c < nc, kk < kn, 0 for kk >= kn;
    // rows of the 16-column tiles past nc are left as they are (they only
    // feed output columns that are never stored)
    const int ne = ((nc + 15) & ~15) * 64;
    for (int e0 = 0; e0 < ne; e0 += 8 * 256) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (e0 + u * 256 >= ne) break; // uniform
            const int e = e0 + u * 256 + (int)threadIdx.x, c = e >> 6, kk = e & 63;
            // unconditional load, padding zeroed by a 0/1 factor (a select
            // would let the compiler branch around the load)
            const double x = src[(size_t)min(c, nc - 1) * cstride + min(kk, kn - 1)];
            const double f = scale ? scale[min(kk, kn - 1)] : 1.0;
            v[u] = x * f * (double)(c < nc && kk < kn);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (e0 + u * 256 >= ne) break;
            const int e = e0 + u * 256 + (int)threadIdx.x;
            sh[(e >> 6) * 65 + (e & 63)] = v[u];
        }
    }
}

// NT: column tiles held in registers (4: 64 columns; 1 for at most 16
// right-hand sides, fewer registers and so more waves per SIMD).
template <int NT = 4>
__device__ __forceinline__ void mm_expand_body(const MmTab &a, int bx, int by,
                                               const double *__restrict__ D,
                                               const double *__restrict__ X, int ldx, double alpha,
                                               double beta, int trans, const double *__restrict__ w,
                                               size_t wstride, double *__restrict__ Y, int ldy)
{
    __shared__ double sh[64 * 65];
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int wv = threadIdx.x >> 6;
    const int rb = bx * 64, r0 = rb + wv * 16;
    const int c0 = by * MM_CG;
    const int N = a.N, b = a.b, nc = min(MM_CG, a.ncols - c0);
    const int nct = min(NT, (nc + 15) >> 4); // live 16-column tiles (uniform)
    // A operand (16 columns c, 4 k): lane holds sh[(16ct + m)*65 + k + g]
    // B operand (4 k x 16 rows of Q^T or D^T): lane holds [k + g][r0 + m]
    // Every global load of a chunk (its B operands, and Y at the start) is
    // issued before the LDS staging of the chunk's A operands, so one round
    // trip serves both; barriers keep the compiler from sinking them.
    d4 acc[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) acc[ct] = d4{0.0, 0.0, 0.0, 0.0};
    // lane holds Y^T[c0 + 16ct + g + 4v][r0 + m]
    double yo[NT][4];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = min(c0 + ct * 16 + g + 4 * v, a.ncols - 1);
            yo[ct][v] = (ct < nct && beta != 0.0) ? Y[(size_t)c * ldy + r0 + m] : 0.0;
        }

    if (D) {
        const int leaf = rb / b, il = r0 - leaf * b;
        const size_t ds_i = trans ? b : 1, ds_j = trans ? 1 : b; // element (row, col) of op(D)
        const double *Dl = D + (size_t)leaf * b * b;
        const double *xl = X + (size_t)c0 * ldx + (size_t)leaf * b;
        for (int j0 = 0; j0 < b; j0 += 32) {
            const int jn = min(32, b - j0);
            double bq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int jj = j0 + min(4 * u, jn - 4) + g;
                bq[u] = Dl[(size_t)(il + m) * ds_i + (size_t)jj * ds_j];
            }
            __syncthreads();
            mm_stage(sh, xl + j0, ldx, nc, jn);
            __syncthreads();
            // rows past the chunk read zeros from sh (and real data from D)
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                if (ct >= nct) continue; // uniform
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc[ct] = mfma64(sh[(16 * ct + m) * 65 + 4 * u + g], bq[u], acc[ct]);
            }
        }
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[ct] *= alpha;
    }
    for (int d = 0; d < a.L; ++d) {
        const int R = a.R[d];
        if (R == 0) continue;
        const int s = rb / a.seglen[d];
        const double *Qd = a.Q[d];
        const double *wd = w + (size_t)c0 * wstride + a.w_off[d] + (size_t)s * R;
        for (int k0 = 0; k0 < R; k0 += 32) {
            const int kn = min(32, R - k0);
            double bq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int kk = k0 + min(4 * u, kn - 4) + g;
                bq[u] = Qd[(size_t)kk * N + r0 + m];
            }
            __syncthreads();
            // apply tables: coefficient = sig (of the block whose U, or V
            // when transposed, lives on segment s) times the stored alpha * t
            mm_stage(sh, wd + k0, wstride, nc, kn,
                     a.sx ? a.S[d] + (size_t)(trans ? (s ^ 1) : s) * R + k0 : nullptr);
            __syncthreads();
            // ranks past the chunk read zeros from sh (and real data from Q)
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                if (ct >= nct) continue; // uniform
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc[ct] = mfma64(sh[(16 * ct + m) * 65 + 4 * u + g], bq[u], acc[ct]);
            }
        }
    }
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = c0 + ct * 16 + g + 4 * v;
            if (c >= a.ncols) continue;
            Y[(size_t)c * ldy + r0 + m] = (beta == 0.0 ? 0.0 : beta * yo[ct][v]) + acc[ct][v];
        }
}



// Y^T (64 columns x 32 rows per wavefront) = beta Y^T + alpha X_leaf^T op(D)^T
// + sum_d W_d^T (sig .* Q_d^T) (sig = 1 with a.nosig), with W rank-major (as
// written by the WT reduction): w[(w_off[d] + s*R + k) * wtc + c].  Four waves per work-group
// cover 128 rows and never wait on each other: the leaf columns (16 per
// chunk, A = X^T, B = op(D)^T, the sum scaled by alpha) and the ranks of all levels
// (16 per chunk, A = W^T, B = sig .* Q^T) form one stream of chunks, the
// next chunk's loads in flight while the current one is multiplied (two
// register buffers, no copies).  Every load is 16 bytes per lane: a lane
// holds two adjacent columns (A tiles ct = 2cp, 2cp+1 are the even and odd
// columns of a 32-column half) and two adjacent rows (row tiles rt = 0, 1
// are the even and odd rows of the wave's 32), or for X of the leaf two
// adjacent leaf columns j (MFMA steps u = 2h, 2h+1 take j = 8h + 2g + {0,1}).
static constexpr int ME_RB = 128;

// The beta term reads C (which may be Y itself); with bmax != null the
// largest |Y| of every (column, 128-row block) goes to bmax[c * (N/128) + bx].
// X, W, Q and D must be 16-byte aligned with even leading dimensions.
template <int TR>
__device__ __forceinline__ void mm_expand2_body(const MmTab &a, int bx, int by,
                                                const double *__restrict__ D,
                                                const double *__restrict__ X, int ldx,
                                                double alpha, double beta,
                                                const double *__restrict__ w, int wtc,
                                                const double *C, int ldc, double *Y, int ldy,
                                                double *__restrict__ bmax)
{
    __shared__ double cmx[4 * MM_CG];
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    // wave-uniform, so that table reads are scalar loads (a vector load
    // would wait for the loads in flight)
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rbk = bx * ME_RB, r0 = rbk + wv * 32;
    const int c0 = by * MM_CG;
    const int N = a.N, b = a.b, nc = min(MM_CG, a.ncols - c0);
    d4 acc[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = d4{0.0, 0.0, 0.0, 0.0};

    // chunks: b/16 of the leaf (when D is given), then 16 ranks at a time
    // over all levels
    const int nlc = D ? b / 16 : 0;
    int nch = 0;
    for (int d = 0; d < a.L; ++d) nch += (a.R[d] + 15) / 16;
    const int ntot = nlc + nch;
    if (ntot > 0) {
        struct Buf
        {
            // level: A[2u + cp] = W^T[c0 + 32cp + 2m + {0,1}][k0 + 4u + g]
            // leaf:  A[4h + q]  = X^T[c0 + 32(q>>1) + 2m + (q&1)][8h + 2g + {0,1}]
            double2 A[8];
            // level, leaf (TR = 0): Bq[u] = B[k(u)][r0 + 2m + {0,1}]
            // leaf (TR = 1): Bq[2rt + h] = op(D)[r0 + 2m + rt][j0 + 8h + 2g + {0,1}]
            double2 Bq[4];
            double sg[4];    // sig of rank k0 + 4u + g (level chunks)
            int kn;          // real ranks in the chunk (0 past the last chunk)
            bool lf;         // leaf chunk
        };
        const int leaf = r0 / b, il = r0 - leaf * b;
        const double *Dl = D ? D + (size_t)leaf * b * b : nullptr;
        const double *Xl = X + (size_t)leaf * b;
        size_t xcol[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            xcol[q] = (size_t)(c0 + min(32 * (q >> 1) + 2 * m + (q & 1), nc - 1)) * ldx;
        // both kinds of chunk load through the same 16 instructions, with
        // addresses selected per chunk (a branch around the loads would make
        // every wait cover the loads in flight)
        auto load = [&](Buf &B, int c) {
            const int cc = min(c, ntot - 1);
            const bool lf = cc < nlc;
            // level chunk (a valid one stands in while the leaf is loaded)
            const int lc = lf ? 0 : cc - nlc;
            int d = 0, cum = 0;
            if (nch > 0)
                while (lc >= cum + (a.R[d] + 15) / 16) cum += (a.R[d++] + 15) / 16;
            const int R = a.R[d], k0 = (lc - cum) * 16, s = r0 / a.seglen[d];
            const double *Wb = w + ((size_t)a.w_off[d] + (size_t)s * R) * wtc + c0 + 2 * m;
            const double *Qb = a.Q[d] + r0 + 2 * m;
            const double *Sb = lf || a.nosig ? Xl : a.S[d] + (size_t)(TR ? (s ^ 1) : s) * R;
            const int j0 = 16 * cc;
            B.kn = c >= ntot ? 0 : lf ? 16 : R - k0;
            B.lf = lf;
            int kc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) kc[u] = min(k0 + 4 * u + g, R - 1);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const double *pl = Wb + (size_t)kc[i >> 1] * wtc + 32 * (i & 1);
                const double *pf = Xl + xcol[i & 3] + j0 + 8 * (i >> 2) + 2 * g;
                B.A[i] = *(const double2 *)(lf ? pf : pl);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double *pl = Qb + (size_t)kc[i] * N;
                const double *pf = TR ? Dl + (size_t)(il + 2 * m + (i >> 1)) * b + j0 + 8 * (i & 1) + 2 * g
                                      : Dl + (size_t)(j0 + 8 * (i >> 1) + 2 * g + (i & 1)) * b + il + 2 * m;
                B.Bq[i] = *(const double2 *)(lf ? pf : pl);
                B.sg[i] = Sb[lf ? 0 : kc[i]];
            }
        };
        auto compute = [&](const Buf &B) {
            if (B.lf) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    double bv[2];
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) {
                        const double2 q2 = TR ? B.Bq[2 * rt + (u >> 1)] : B.Bq[u];
                        bv[rt] = TR ? ((u & 1) ? q2.y : q2.x) : (rt ? q2.y : q2.x);
                    }
#pragma unroll
                    for (int ct = 0; ct < 4; ++ct) {
                        const double2 a2 = B.A[4 * (u >> 1) + ct];
                        const double av = (u & 1) ? a2.y : a2.x;
#pragma unroll
                        for (int rt = 0; rt < 2; ++rt) acc[rt][ct] = mfma64(av, bv[rt], acc[rt][ct]);
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double f = 4 * u + g < B.kn ? (a.nosig ? 1.0 : B.sg[u]) : 0.0;
                    const double bv[2] = {B.Bq[u].x * f, B.Bq[u].y * f};
#pragma unroll
                    for (int ct = 0; ct < 4; ++ct) {
                        const double2 a2 = B.A[2 * u + (ct >> 1)];
                        const double av = (ct & 1) ? a2.y : a2.x;
#pragma unroll
                        for (int rt = 0; rt < 2; ++rt) acc[rt][ct] = mfma64(av, bv[rt], acc[rt][ct]);
                    }
                }
            }
        };
        // the leaf term (chunks 0 .. nlc-1, nlc even) is scaled by alpha
        // once it is complete; W of the levels carries alpha already
        auto scale = [&]() {
#pragma unroll
            for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) acc[rt][ct] *= alpha;
        };
        Buf bA, bB;
        load(bA, 0);
#pragma unroll 1
        for (int c = 0; c < ntot; c += 2) {
            load(bB, c + 1);
            if (nlc > 0 && c == nlc) scale();
            compute(bA);
            load(bA, c + 2);
            compute(bB);
        }
        if (nlc > 0 && nch == 0) scale();
    }

    // lane holds Y^T[c0 + 32(ct>>1) + 2(g + 4v) + (ct&1)][r0 + 2m + rt]
    double yo[2][4][4];
    if (beta != 0.0) {
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int c = min(c0 + 32 * (ct >> 1) + 2 * (g + 4 * v) + (ct & 1), a.ncols - 1);
                    yo[rt][ct][v] = C[(size_t)c * ldc + r0 + 2 * m + rt];
                }
    }
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = c0 + 32 * (ct >> 1) + 2 * (g + 4 * v) + (ct & 1);
                acc[rt][ct][v] += beta == 0.0 ? 0.0 : beta * yo[rt][ct][v];
                if (c < a.ncols) Y[(size_t)c * ldy + r0 + 2 * m + rt] = acc[rt][ct][v];
            }
    if (bmax) {
        // column maxima: over the two row tiles, the 16 lanes of a group
        // (rows), then the four waves
        double cm[4][4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double x = fmax(fabs(acc[0][ct][v]), fabs(acc[1][ct][v]));
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) x = fmax(x, __shfl_xor(x, o, 64));
                cm[ct][v] = x;
            }
        if (m == 0)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    cmx[wv * 64 + 32 * (ct >> 1) + 2 * (g + 4 * v) + (ct & 1)] = cm[ct][v];
        __syncthreads();
        if (threadIdx.x < MM_CG && c0 + (int)threadIdx.x < a.ncols) {
            const int c = threadIdx.x;
            const double x = fmax(fmax(cmx[c], cmx[64 + c]), fmax(cmx[128 + c], cmx[192 + c]));
            bmax[(size_t)(c0 + c) * (N / ME_RB) + bx] = x;
        }
    }
}

// Expansion, third form: as mm_expand2_body on 256 rows x 64 columns per
// work-group (eight waves of 32 rows), with W^T and sig of the levels whose
// segments hold the whole work-group ("shared" levels, segments of 256 rows
// or more, a prefix of the levels) staged in LDS once, behind the one
// barrier, so that every wave takes them from LDS instead of each reading
// them from global memory.  The stream then runs in three phases, each
// with its own two register buffers: the leaf (A = X^T, B = op(D)^T, from
// global memory), the shared levels (A from LDS, B = Q from global memory),
// the other levels (as in the second form).  bmax covers the work-group's
// two 128-row blocks.  A staged row is 66 doubles (16-byte aligned), so a
// lane reads its two adjacent columns as one 16-byte LDS read, and the 16
// lanes of a group read 64 consecutive dwords (every bank once).
static constexpr int ME3_RB = 256, ME3_NW = 8, ME3_CS = 66;

// LDS doubles of the third form for `rows` staged ranks (W^T rows, sig,
// row offsets and sig addresses).
__host__ __device__ inline size_t me3_lds(int rows) { return (size_t)rows * (ME3_CS + 3); }

template <int TR>
__device__ __forceinline__ void mm_expand3_body(const MmTab &a, int bx, int by,
                                                const double *__restrict__ D,
                                                const double *__restrict__ X, int ldx,
                                                double alpha, double beta,
                                                const double *__restrict__ w, int wtc,
                                                const double *C, int ldc, double *Y, int ldy,
                                                double *__restrict__ bmax, double *lds, int NC = -1)
{
    __shared__ double cmx[ME3_NW * MM_CG];
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int rb0 = bx * ME3_RB, r0 = rb0 + wv * 32;
    const int c0 = by * MM_CG;
    if (NC < 0) NC = a.ncols; // columns (a split product gives each part its own)
    const int N = a.N, b = a.b, nc = min(MM_CG, NC - c0);

    // shared levels [0, dsh) and their staged rows
    int dsh = 0, rows = 0;
    while (dsh < a.L && a.seglen[dsh] >= ME3_RB) rows += a.R[dsh++];
    double *Ws = lds;                    // [rows][ME3_CS]: W^T[c0 + c][k] at Ws[k * CS + c]
    double *Ss = lds + (size_t)rows * ME3_CS; // sig of the staged ranks
    // element offset in w and sig address of each staged row; the level
    // loop is uniform, so the table entries come from scalar loads (a
    // per-lane level search would read the kernel arguments through
    // dependent vector loads), and the loop itself loads nothing
    long long *ro = (long long *)(Ss + rows);
    const double **sp = (const double **)(ro + rows);
    {
        int row0 = 0;
        for (int d = 0; d < dsh; ++d) {
            const int R = a.R[d], s = rb0 / a.seglen[d];
            const double *Sd = a.S[d] + (size_t)(TR ? (s ^ 1) : s) * R;
            for (int k = threadIdx.x; k < R; k += 64 * ME3_NW) {
                ro[row0 + k] = ((long long)a.w_off[d] + (long long)s * R + k) * wtc + c0;
                sp[row0 + k] = Sd + k;
            }
            row0 += R;
        }
    }
    __syncthreads();
    // the sig of row tid goes out with the first batch (a valid dummy
    // address when there is nothing to stage; the table is not read then)
    const double *spt = w;
    if (rows > 0 && !a.nosig) spt = sp[min((int)threadIdx.x, rows - 1)];
    const double sgl = *spt;
    // eight pairs per thread in flight at once (a loop of single loads and
    // stores would wait on memory once per pair); each work-group starts at
    // another row, so that the many work-groups sharing a top-level segment
    // do not all ask for the same lines at the same moment
    const int ne = rows * 32, rot = rows > 0 ? (bx * 37 % rows) * 32 : 0;
    for (int e0 = 0; e0 < ne; e0 += 8 * 64 * ME3_NW) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int e = min(e0 + u * 64 * ME3_NW + (int)threadIdx.x, ne - 1) + rot;
            e -= e >= ne ? ne : 0;
            v[u] = *(const double2 *)(w + ro[e >> 5] + 2 * (e & 31));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int e = e0 + u * 64 * ME3_NW + (int)threadIdx.x;
            if (e < ne) {
                e += rot;
                e -= e >= ne ? ne : 0;
                *(double2 *)(Ws + (e >> 5) * ME3_CS + 2 * (e & 31)) = v[u];
            }
        }
    }
    for (int r = threadIdx.x; r < rows; r += 64 * ME3_NW)
        Ss[r] = a.nosig ? 1.0 : (r == (int)threadIdx.x ? sgl : *sp[r]);
    __syncthreads();

    d4 acc[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = d4{0.0, 0.0, 0.0, 0.0};

    // phase 1: the leaf, b/16 chunks of 16 leaf columns j (steps u = 2h, 2h+1
    // take j = 8h + 2g + {0,1}), A = X^T pairs, B = op(D)^T
    if (D) {
        struct LBuf
        {
            double2 A[8];  // A[4h + q] = X^T[c0 + 32(q>>1) + 2m + (q&1)][j0 + 8h + 2g + {0,1}]
            double2 Bq[4]; // TR = 0: Bq[u] = op(D)[r0 + 2m + {0,1}][j(u)]; TR = 1: Bq[2rt + h]
        };
        const int leaf = r0 / b, il = r0 - leaf * b;
        const double *Dl = D + (size_t)leaf * b * b;
        const double *Xl = X + (size_t)leaf * b;
        size_t xcol[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            xcol[q] = (size_t)(c0 + min(32 * (q >> 1) + 2 * m + (q & 1), nc - 1)) * ldx;
        const int nlc = b / 16;
        auto load = [&](LBuf &B, int c) {
            const int j0 = 16 * min(c, nlc - 1);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                B.A[i] = *(const double2 *)(Xl + xcol[i & 3] + j0 + 8 * (i >> 2) + 2 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                B.Bq[i] = *(const double2 *)(TR ? Dl + (size_t)(il + 2 * m + (i >> 1)) * b + j0 + 8 * (i & 1) + 2 * g
                                                : Dl + (size_t)(j0 + 8 * (i >> 1) + 2 * g + (i & 1)) * b + il + 2 * m);
        };
        auto compute = [&](const LBuf &B, int c) {
            if (c >= nlc) return;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double bv[2];
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) {
                    const double2 q2 = TR ? B.Bq[2 * rt + (u >> 1)] : B.Bq[u];
                    bv[rt] = TR ? ((u & 1) ? q2.y : q2.x) : (rt ? q2.y : q2.x);
                }
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    const double2 a2 = B.A[4 * (u >> 1) + ct];
                    const double av = (u & 1) ? a2.y : a2.x;
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) acc[rt][ct] = mfma64(av, bv[rt], acc[rt][ct]);
                }
            }
        };
        LBuf bA, bB;
        load(bA, 0);
#pragma unroll 1
        for (int c = 0; c < nlc; c += 2) {
            load(bB, c + 1);
            compute(bA, c);
            load(bA, c + 2);
            compute(bB, c + 1);
        }
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[rt][ct] *= alpha;
    }

    // phase 2: the shared levels, 16 ranks per chunk; A (W^T) and sig from LDS
    int nsc = 0;
    for (int d = 0; d < dsh; ++d) nsc += (a.R[d] + 15) / 16;
    if (nsc > 0) {
        struct SBuf
        {
            double2 Bq[4]; // Q[k(u)][r0 + 2m + {0,1}]
            int row0, kn;  // first staged row of the chunk, real ranks in it
        };
        auto load = [&](SBuf &B, int c) {
            const int cc = min(c, nsc - 1);
            int d = 0, cum = 0, pre = 0;
            while (cc >= cum + (a.R[d] + 15) / 16) {
                cum += (a.R[d] + 15) / 16;
                pre += a.R[d++];
            }
            const int R = a.R[d], k0 = (cc - cum) * 16;
            B.row0 = pre + k0;
            B.kn = c < nsc ? R - k0 : 0;
            const double *Qb = a.Q[d] + r0 + 2 * m;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                B.Bq[u] = *(const double2 *)(Qb + (size_t)min(k0 + 4 * u + g, R - 1) * N);
        };
        auto compute = [&](const SBuf &B) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int kk = 4 * u + g, row = B.row0 + min(kk, max(B.kn - 1, 0));
                const double f = kk < B.kn ? Ss[row] : 0.0;
                const double bv[2] = {B.Bq[u].x * f, B.Bq[u].y * f};
                const double *wr = Ws + (size_t)row * ME3_CS + 2 * m;
#pragma unroll
                for (int cp = 0; cp < 2; ++cp) {
                    const double2 a01 = *(const double2 *)(wr + 32 * cp);
                    const double a0 = a01.x, a1 = a01.y;
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) {
                        acc[rt][2 * cp] = mfma64(a0, bv[rt], acc[rt][2 * cp]);
                        acc[rt][2 * cp + 1] = mfma64(a1, bv[rt], acc[rt][2 * cp + 1]);
                    }
                }
            }
        };
        SBuf bA, bB;
        load(bA, 0);
#pragma unroll 1
        for (int c = 0; c < nsc; c += 2) {
            load(bB, c + 1);
            compute(bA);
            load(bA, c + 2);
            compute(bB);
        }
    }

    // phase 3: the other levels, as in the second form
    int nch = 0;
    for (int d = dsh; d < a.L; ++d) nch += (a.R[d] + 15) / 16;
    if (nch > 0) {
        struct Buf
        {
            double2 A[8];  // A[2u + cp] = W^T[c0 + 32cp + 2m + {0,1}][k0 + 4u + g]
            double2 Bq[4]; // Q[k(u)][r0 + 2m + {0,1}]
            double sg[4];
            int kn;
        };
        auto load = [&](Buf &B, int c) {
            const int cc = min(c, nch - 1);
            int d = dsh, cum = 0;
            while (cc >= cum + (a.R[d] + 15) / 16) cum += (a.R[d++] + 15) / 16;
            const int R = a.R[d], k0 = (cc - cum) * 16, s = r0 / a.seglen[d];
            const double *Wb = w + ((size_t)a.w_off[d] + (size_t)s * R) * wtc + c0 + 2 * m;
            const double *Qb = a.Q[d] + r0 + 2 * m;
            const double *Sb = a.nosig ? Qb : a.S[d] + (size_t)(TR ? (s ^ 1) : s) * R;
            B.kn = c < nch ? R - k0 : 0;
            int kc[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) kc[u] = min(k0 + 4 * u + g, R - 1);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                B.A[i] = *(const double2 *)(Wb + (size_t)kc[i >> 1] * wtc + 32 * (i & 1));
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                B.Bq[i] = *(const double2 *)(Qb + (size_t)kc[i] * N);
                B.sg[i] = Sb[a.nosig ? 0 : kc[i]];
            }
        };
        auto compute = [&](const Buf &B) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double f = 4 * u + g < B.kn ? (a.nosig ? 1.0 : B.sg[u]) : 0.0;
                const double bv[2] = {B.Bq[u].x * f, B.Bq[u].y * f};
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    const double2 a2 = B.A[2 * u + (ct >> 1)];
                    const double av = (ct & 1) ? a2.y : a2.x;
#pragma unroll
                    for (int rt = 0; rt < 2; ++rt) acc[rt][ct] = mfma64(av, bv[rt], acc[rt][ct]);
                }
            }
        };
        Buf bA, bB;
        load(bA, 0);
#pragma unroll 1
        for (int c = 0; c < nch; c += 2) {
            load(bB, c + 1);
            compute(bA);
            load(bA, c + 2);
            compute(bB);
        }
    }

    // lane holds Y^T[c0 + 32(ct>>1) + 2(g + 4v) + (ct&1)][r0 + 2m + rt]
    double yo[2][4][4];
    if (beta != 0.0) {
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int c = min(c0 + 32 * (ct >> 1) + 2 * (g + 4 * v) + (ct & 1), NC - 1);
                    yo[rt][ct][v] = C[(size_t)c * ldc + r0 + 2 * m + rt];
                }
    }
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = c0 + 32 * (ct >> 1) + 2 * (g + 4 * v) + (ct & 1);
                acc[rt][ct][v] += beta == 0.0 ? 0.0 : beta * yo[rt][ct][v];
                if (c < NC) Y[(size_t)c * ldy + r0 + 2 * m + rt] = acc[rt][ct][v];
            }
    if (bmax) {
        // column maxima per 128-row block: the two row tiles, the 16 lanes of
        // a group, then the four waves of the block
        double cm[4][4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double x = fmax(fabs(acc[0][ct][v]), fabs(acc[1][ct][v]));
#pragma unroll
                for (int o = 1; o < 16; o <<= 1) x = fmax(x, __shfl_xor(x, o, 64));
                cm[ct][v] = x;
            }
        if (m == 0)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    cmx[wv * 64 + 32 * (ct >> 1) + 2 * (g + 4 * v) + (ct & 1)] = cm[ct][v];
        __syncthreads();
        if (threadIdx.x < 2 * MM_CG) {
            const int c = threadIdx.x & 63, h = threadIdx.x >> 6;
            if (c0 + c < NC) {
                const double *q = cmx + h * 4 * 64 + c;
                const double x = fmax(fmax(q[0], q[64]), fmax(q[128], q[192]));
                bmax[(size_t)(c0 + c) * (N / ME_RB) + 2 * bx + h] = x;
            }
        }
    }
}

// Table for levels [d0, d1) with operands P (reduced), Q (expanded) and S.
static inline void mm_tab(const dplrh_plan *p, int d0, int d1, const double *const *P,
                   const double *const *Q, const double *const *S, int ncols, int w_local,
                   MmTab *t, int tmax = MM_TMAX)
{
    memset(t, 0, sizeof(*t));
    t->L = d1 - d0;
    t->N = p->N;
    t->b = p->b;
    t->ncols = ncols;
    t->ncg = (ncols + MM_CG - 1) / MM_CG;
    t->sx = !w_local;
    int task = 0, part = 0, fin = 0, woff = 0;
    for (int i = 0; i < t->L; ++i) {
        const int d = d0 + i;
        const int seg = p->N >> (d + 1), nseg = 2 << d, R = p->R[d];
        const int T = seg < tmax ? seg : tmax;
        const int tps = seg / T;
        const int nkg = (R + MM_KG - 1) / MM_KG;
        t->R[i] = R;
        t->seglen[i] = seg;
        t->T[i] = T;
        t->tps[i] = tps;
        t->nkg[i] = nkg;
        t->task_off[i] = task;
        task += nkg * t->ncg * nseg * tps;
        t->part_off[i] = part;
        if (tps > 1) part += nkg * t->ncg * nseg * tps;
        t->fin_off[i] = fin;
        if (tps > 1) fin += nkg * t->ncg * nseg;
        t->w_off[i] = w_local ? woff : (int)p->soff[d];
        woff += nseg * R;
        t->P[i] = P[i];
        t->Q[i] = Q[i];
        t->S[i] = S[i];
    }
    t->task_off[t->L] = task;
    t->part_off[t->L] = part;
    t->fin_off[t->L] = fin;
    t->w_off[t->L] = w_local ? woff : (int)p->soff[d1];
}
