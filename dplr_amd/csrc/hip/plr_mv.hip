// PLR matvec and its column-batched form: y := alpha*op(A)*x + beta*y.
//
// The reference walks the tree depth-first and issues two GEMVs per
// low-rank block and one per leaf (src/dplrpmv.c:19-180).  Here the whole
// tree is two launches over the level-wise arrays:
//
//   mv_reduce  one wavefront per (level, rank group, segment, tile) reads a
//              tile of P = V (or U for A^T) and x, reduces the tile's
//              contribution to t = P|_s^T x|_s across the wavefront, and the
//              last tile to arrive for a segment sums the tiles' partials in
//              a fixed order and writes w_{s^1} = alpha * sig .* t_s.
//   mv_expand  one thread per row and rank quarter: leaf product plus
//              sum over levels of Q|_s w_s (Q = U, or V for A^T).
//
// Both kernels take a column count: blockIdx.y selects the column, so the
// same code serves the multi-RHS product for any shape (each column reads
// the factors once).
#include <cstring>
#include <mutex>

#include "plr_common.h"
#include "dplrx.h"

static constexpr int KG = 16; // rank group; matches plr_plan.hip

// Sum over the 64 lanes of a wavefront; every lane receives the total.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_mv_reduce(LevelTab a, const double *__restrict__ x,
                                                   int ldx, double alpha, int trans, int raw,
                                                   double *__restrict__ part, size_t part_stride,
                                                   int *__restrict__ cnt, size_t cnt_stride,
                                                   double *__restrict__ w, size_t w_stride)
{
    const int lane = threadIdx.x & 63;
    const int task = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (task >= a.task_off[a.L]) return;
    const int col = blockIdx.y;
    x += (size_t)col * ldx;
    part += (size_t)col * part_stride;
    cnt += (size_t)col * cnt_stride;
    w += (size_t)col * w_stride;

    int d = 0;
    while (task >= a.task_off[d + 1]) ++d;
    const int loc = task - a.task_off[d];
    const int nkg = a.nkg[d], tps = a.tps[d];
    const int kg = loc % nkg;
    const int st = loc / nkg; // segment * tps + tile
    const int s = st / tps, tile = st - s * tps;
    const int R = a.R[d], k0 = kg * KG;
    const int kn = min(KG, R - k0);
    const int N = a.N, seg = a.seglen[d], T = a.T[d];
    const int rbeg = s * seg + tile * T;
    const int rend = min(rbeg + T, (s + 1) * seg);
    // Columns past the rank group's end load a duplicate of its last column:
    // every load is unconditional (no branch, no wait per load) and the
    // duplicates' sums are never written.
    const double *__restrict__ P = a.P[d] + (size_t)k0 * N;
    size_t coff[KG];
#pragma unroll
    for (int kk = 0; kk < KG; ++kk) coff[kk] = (size_t)(kk < kn ? kk : kn - 1) * N;

    double acc[KG];
#pragma unroll
    for (int kk = 0; kk < KG; ++kk) acc[kk] = 0.0;
    for (int i = rbeg + lane; i < rend; i += 64) {
        double pv[KG];
#pragma unroll
        for (int kk = 0; kk < KG; ++kk) pv[kk] = __builtin_nontemporal_load(P + coff[kk] + i);
        const double xv = x[i];
#pragma unroll
        for (int kk = 0; kk < KG; ++kk) acc[kk] += pv[kk] * xv;
    }
#pragma unroll
    for (int kk = 0; kk < KG; ++kk) acc[kk] = wave_sum(acc[kk]);

    // lane kk owns rank column k0 + kk from here on
    double mine = 0.0;
#pragma unroll
    for (int kk = 0; kk < KG; ++kk)
        if (lane == kk) mine = acc[kk];

    // raw: t_s itself, stored at segment s.  Otherwise the coefficients
    // alpha*sig.*t_s that act on the partner segment s^1.
    const int sp = raw ? s : (s ^ 1);
    const int sidx = trans ? s : (s ^ 1); // whose sig scales them
    const double *S = raw ? nullptr : a.S[d] + (size_t)sidx * R;
    double *wd = w + a.w_off[d] + (size_t)sp * R;

    if (tps == 1) {
        if (lane < kn) wd[k0 + lane] = raw ? mine : alpha * S[k0 + lane] * mine;
        return;
    }
    double *pd = part + a.part_off[d];
    if (lane < kn) pd[(size_t)st * R + k0 + lane] = mine;
    __threadfence();
    int old = 0;
    int *c = cnt + a.cnt_off[d] + s * nkg + kg;
    if (lane == 0) old = atomicAdd(c, 1);
    old = __shfl(old, 0, 64);
    if (old != tps - 1) return;
    // Last tile of this segment: every other tile's partial is visible.  All
    // 64 lanes sum: lane = rank (16) x tile phase (4), four independent
    // accumulators each, so the loads overlap; then the phases combine.
    __threadfence();
    const int kk = lane & 15, ph = lane >> 4;
    const int kc = k0 + (kk < kn ? kk : kn - 1);
    const double *q = pd + (size_t)s * tps * R + kc;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    for (int j = ph; j < tps; j += 16) {
        const int j1 = min(j + 4, tps - 1), j2 = min(j + 8, tps - 1), j3 = min(j + 12, tps - 1);
        const double v0 = q[(size_t)j * R], v1 = q[(size_t)j1 * R];
        const double v2 = q[(size_t)j2 * R], v3 = q[(size_t)j3 * R];
        t0 += v0;
        t1 += (j + 4 < tps) ? v1 : 0.0;
        t2 += (j + 8 < tps) ? v2 : 0.0;
        t3 += (j + 12 < tps) ? v3 : 0.0;
    }
    double t = (t0 + t1) + (t2 + t3);
    t += __shfl_xor(t, 16, 64);
    t += __shfl_xor(t, 32, 64);
    if (ph == 0 && kk < kn) wd[k0 + kk] = raw ? t : alpha * S[k0 + kk] * t;
    if (lane == 0) atomicExch(c, 0);
}

__global__ __launch_bounds__(256) void k_mv_expand(LevelTab a, const double *__restrict__ D,
                                                   const double *__restrict__ x, int ldx,
                                                   double alpha, double beta, int trans,
                                                   const double *__restrict__ w, size_t w_stride,
                                                   double *__restrict__ y, int ldy)
{
    __shared__ double red[3][64];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = blockIdx.y;
    x += (size_t)col * ldx;
    y += (size_t)col * ldy;
    w += (size_t)col * w_stride;
    const int N = a.N, b = a.b;
    const int row0 = blockIdx.x * 64;
    const int i = row0 + lane;
    double acc = 0.0;
    if (i < N) {
        if (D) { // leaf term (skipped when the caller passes no leaves)
            const int leaf = i / b, il = i - leaf * b;
            const double *__restrict__ Dl = D + (size_t)leaf * b * b;
            const double *__restrict__ xl = x + (size_t)leaf * b;
            double sl = 0.0;
            if (!trans) {
                for (int j = g; j < b; j += 4) sl += Dl[(size_t)j * b + il] * xl[j];
            } else {
                for (int j = g; j < b; j += 4) sl += Dl[(size_t)il * b + j] * xl[j];
            }
            acc = alpha * sl;
        }
        for (int d = 0; d < a.L; ++d) {
            const int R = a.R[d], seg = a.seglen[d];
            const double *__restrict__ Q = a.Q[d];
            if ((seg & 63) == 0) {
                // the wavefront's 64 rows share one segment: uniform coefficients
                const int s = __builtin_amdgcn_readfirstlane(row0 / seg);
                const double *__restrict__ wd = w + a.w_off[d] + (size_t)s * R;
#pragma unroll 4
                for (int k = g; k < R; k += 4) acc += Q[(size_t)k * N + i] * wd[k];
            } else {
                const int s = i / seg;
                const double *__restrict__ wd = w + a.w_off[d] + (size_t)s * R;
                for (int k = g; k < R; k += 4) acc += Q[(size_t)k * N + i] * wd[k];
            }
        }
    }
    if (g > 0) red[g - 1][lane] = acc;
    __syncthreads();
    if (g == 0 && i < N) {
        const double r = acc + red[0][lane] + red[1][lane] + red[2][lane];
        y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + r;
    }
}

// ---------------------------------------------------------------------------
// Fast path for segments that are whole multiples of 64 rows (leaf order a
// multiple of 64).  Reduction: one wavefront per (level, 16 ranks, block of
// rows); lane l streams rows 4l..4l+3 of each 256-row slab (two 16-byte
// loads per rank column, all issued before use).  The 64 lane sums of each
// rank meet once in LDS; a lane quarter is 64 consecutive rows, so segments
// of 64, 128 or >= 256 rows come out of the same reduction.  Tiles of long
// segments write partials that k_mv_finish sums (no fences, no counters).

typedef double d2v __attribute__((ext_vector_type(2)));

struct Mv2Tab
{
    int L, N, nb, nch; // reduction batches of KG ranks, expansion chunks of 32 ranks
    int trans;
    int R[PLR_MAXL], seg[PLR_MAXL], TW[PLR_MAXL], tps[PLR_MAXL], nkg[PLR_MAXL];
    int task_off[PLR_MAXL + 1], part_off[PLR_MAXL + 1], fin_off[PLR_MAXL + 1];
    int w_off[PLR_MAXL + 1];
    const double *P[PLR_MAXL], *Q[PLR_MAXL], *S[PLR_MAXL];
};

__device__ __forceinline__ double mv_coef(double t, int raw, double alpha, const double *S,
                                          int sidx, int R, int k)
{
    return raw ? t : alpha * S[(size_t)sidx * R + k] * t;
}

// Reduction t = P|_s^T x|_s for every level at once: a work-group owns 128
// rows (lane l holds rows 2l, 2l+1, so x is read once for all levels) and
// its four waves take the level-wise batches of KG columns round-robin.  The
// 64-lane sums go through a per-wave LDS transpose; segments of 64 or 128
// rows are finished here, longer ones leave one partial per 128-row chunk
// for k_mv_finish.  Every lane issues its KG 16-byte loads back to back.
// The sum over segment s times the sig of its block goes to w at segment
// s ^ 1 (the rows it is expanded into); the expansion applies alpha.
__global__ __launch_bounds__(256) void k_mv_reduce2(Mv2Tab a, const double *__restrict__ x,
                                                    int ldx, double *__restrict__ part,
                                                    size_t pstride, double *__restrict__ w,
                                                    size_t wstride)
{
    __shared__ double red[4][KG][65];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = blockIdx.y;
    const int N = a.N, row0 = blockIdx.x * 128, i = row0 + 2 * lane;
    x += (size_t)col * ldx;
    part += (size_t)col * pstride;
    w += (size_t)col * wstride;
    const d2v xv = *(const d2v *)(x + i);
    // batch bi -> (level, first column) by a walk of the level table
    auto locate = [&](int bi, int &dl, int &kl) {
        int bd = 0;
        dl = 0;
        while (bi >= bd + a.nkg[dl]) bd += a.nkg[dl++];
        kl = (bi - bd) * KG;
    };
    auto load = [&](d2v *v, int dl, int kl) {
        const int kn = min(KG, a.R[dl] - kl);
        const double *__restrict__ P = a.P[dl] + (size_t)kl * N + i;
#pragma unroll
        for (int kk = 0; kk < KG; ++kk)
            v[kk] = __builtin_nontemporal_load((const d2v *)(P + (size_t)(kk < kn ? kk : kn - 1) * N));
    };
    auto reduce = [&](const d2v *v, int dl, int kl) {
        const int R = a.R[dl], kn = min(KG, R - kl);
#pragma unroll
        for (int kk = 0; kk < KG; ++kk) red[g][kk][lane] = v[kk].x * xv.x + v[kk].y * xv.y;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // lane (kk, quarter) sums the 16 lanes (32 rows) of one quarter
        const int kk = lane & 15, qt = lane >> 4;
        double sq = 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) sq += red[g][kk][qt * 16 + j];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int kc = kl + kk, seg = a.seg[dl];
        const bool ok = kk < kn;
        sq += __shfl_xor(sq, 16, 64); // 64 rows
        if (seg == 64) {
            const int s = row0 / 64 + (qt >> 1);
            if (ok && (qt & 1) == 0)
                w[a.w_off[dl] + (size_t)(s ^ 1) * R + kc] = sq * a.S[dl][(size_t)(a.trans ? s : s ^ 1) * R + kc];
            return;
        }
        sq += __shfl_xor(sq, 32, 64); // 128 rows
        if (!ok || qt != 0) return;
        const int s = row0 / seg;
        if (seg == 128) {
            w[a.w_off[dl] + (size_t)(s ^ 1) * R + kc] = sq * a.S[dl][(size_t)(a.trans ? s : s ^ 1) * R + kc];
        } else {
            const int tile = (row0 - s * seg) / 128;
            part[a.part_off[dl] + ((size_t)s * a.tps[dl] + tile) * R + kc] = sq;
        }
    };
    // Software pipeline: the loads of the wave's next batch are issued before
    // the current batch is reduced.  Two buffers alternate (a register copy
    // would wait for the prefetch) and no load sits under a branch that
    // rejoins before its use (the join would wait for every load).
    const int cnt = g < a.nb ? (a.nb - g + 3) / 4 : 0;
    if (cnt == 0) return;
    d2v b0[KG], b1[KG];
    int d0, k0, d1, k1;
    locate(g, d0, k0);
    load(b0, d0, k0);
    for (int it = 0;; it += 2) {
        if (it + 1 >= cnt) {
            reduce(b0, d0, k0);
            break;
        }
        locate(g + 4 * (it + 1), d1, k1);
        load(b1, d1, k1);
        reduce(b0, d0, k0);
        if (it + 2 >= cnt) {
            reduce(b1, d1, k1);
            break;
        }
        locate(g + 4 * (it + 2), d0, k0);
        load(b0, d0, k0);
        reduce(b1, d1, k1);
    }
}

// Sum the tile partials of long segments (fixed order) and scale the sum by
// the sig of its block: a work-group covers 32 consecutive (level, segment,
// rank) entries with eight tile phases each (eight loads in flight per
// thread), combined through LDS.
__global__ __launch_bounds__(256) void k_mv_finish(Mv2Tab a, const double *__restrict__ part,
                                                   size_t pstride, double *__restrict__ w,
                                                   size_t wstride)
{
    __shared__ double red[8][32];
    const int el = threadIdx.x & 31, ph = threadIdx.x >> 5;
    const int e = blockIdx.x * 32 + el;
    const int col = blockIdx.y;
    const bool ok = e < a.fin_off[a.L];
    int d = 0;
    if (ok)
        while (e >= a.fin_off[d + 1]) ++d;
    const int loc = ok ? e - a.fin_off[d] : 0, R = a.R[d], tps = a.tps[d];
    const int s = loc / (R ? R : 1), k = loc - s * R;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ok) {
        const double *pp = part + (size_t)col * pstride + a.part_off[d] + (size_t)s * tps * R + k;
        for (int j0 = ph; j0 < tps; j0 += 64) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int j = j0 + 8 * u;
                const double v = pp[(size_t)min(j, tps - 1) * R];
                acc[u] += j < tps ? v : 0.0;
            }
        }
    }
    red[ph][el] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (ph != 0 || !ok) return;
    const double t = ((red[0][el] + red[1][el]) + (red[2][el] + red[3][el])) +
                     ((red[4][el] + red[5][el]) + (red[6][el] + red[7][el]));
    w[(size_t)col * wstride + a.w_off[d] + (size_t)(s ^ 1) * R + k] =
        t * a.S[d][(size_t)(a.trans ? s : s ^ 1) * R + k];
}

// Expansion: work-group = 128 rows x 4 rank quarters; lane l holds rows
// 2l, 2l+1 (16-byte loads), each rank quarter's loads issued in batches of 8.
__global__ __launch_bounds__(256) void k_mv_expand2(Mv2Tab a, int b, const double *__restrict__ D,
                                                    const double *__restrict__ x, int ldx,
                                                    double alpha, double beta, int trans,
                                                    const double *__restrict__ w, size_t wstride,
                                                    double *__restrict__ y, int ldy)
{
    __shared__ double2 red[3][64];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = blockIdx.y;
    x += (size_t)col * ldx;
    y += (size_t)col * ldy;
    w += (size_t)col * wstride;
    const int N = a.N;
    const int row0 = blockIdx.x * 128;
    const int i = row0 + 2 * lane;
    double2 lr = {0.0, 0.0}; // sum over the levels, before alpha
    // The levels are walked as chunks of 32 ranks (wave g takes ranks
    // k0 + g + 4u), with the buffers in rotation (a register copy, or a load
    // under a branch that rejoins, would wait for the prefetch).
    auto locate = [&](int ci, int &dl, int &kl) {
        int cb = 0;
        dl = 0;
        while (ci >= cb + (a.R[dl] + 31) / 32) cb += (a.R[dl++] + 31) / 32;
        kl = (ci - cb) * 32;
    };
    // (w of segment s already carries the sig of the block whose U, V when
    // transposed, lives on s; alpha is applied at the end)
    auto load = [&](int dl, int kl, d2v *qv, double *wr) {
        const int R = a.R[dl], seg = a.seg[dl];
        const bool uni = (seg & 127) == 0;
        const int s = uni ? __builtin_amdgcn_readfirstlane(row0 / seg) : i / seg;
        const double *__restrict__ Q = a.Q[dl];
        const double *__restrict__ wd = w + a.w_off[dl] + (size_t)s * R;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = min(kl + g + 4 * u, R - 1);
            qv[u] = __builtin_nontemporal_load((const d2v *)(Q + (size_t)k * N + i));
            wr[u] = wd[k];
        }
    };
    auto comp = [&](int dl, int kl, const d2v *qv, const double *wr) {
        const int R = a.R[dl];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double c = wr[u] * (double)(kl + g + 4 * u < R);
            lr.x += qv[u].x * c;
            lr.y += qv[u].y * c;
        }
    };
    // three chunk buffers in rotation; the first two chunks' loads go out
    // before the leaf product, so they are in flight while the leaf waits on
    // its own (chunk 1 is clamped to chunk 0 when there is only one)
    d2v q0[8], q1[8], q2[8];
    double w0[8], w1[8], w2[8];
    int d0 = 0, k0 = 0, d1 = 0, k1 = 0, d2, k2;
    if (a.nch > 0) {
        locate(0, d0, k0);
        load(d0, k0, q0, w0);
        locate(min(1, a.nch - 1), d1, k1);
        load(d1, k1, q1, w1);
    }
    double2 acc = {0.0, 0.0};
    if (D) {
        const int leaf = i / b, il = i - leaf * b;
        const double *__restrict__ Dl = D + (size_t)leaf * b * b;
        const double *__restrict__ xl = x + (size_t)leaf * b;
        double2 sl = {0.0, 0.0};
        if (!trans) {
            // sixteen leaf columns per wave in flight at once
            for (int j0 = g; j0 < b; j0 += 64) {
                double2 dv[16];
                double xv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int j = min(j0 + 4 * u, b - 1);
                    dv[u] = *(const double2 *)(Dl + (size_t)j * b + il);
                    xv[u] = xl[j];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const double xu = j0 + 4 * u < b ? xv[u] : 0.0;
                    sl.x += dv[u].x * xu;
                    sl.y += dv[u].y * xu;
                }
            }
        } else {
            for (int j = g; j < b; j += 4) {
                sl.x += Dl[(size_t)il * b + j] * xl[j];
                sl.y += Dl[(size_t)(il + 1) * b + j] * xl[j];
            }
        }
        acc.x = alpha * sl.x;
        acc.y = alpha * sl.y;
    }
    // (the loads of chunk it + 2 are issued before chunk it is used; a
    // branch either exits the loop or holds no load)
    if (a.nch > 0) {
        for (int it = 0;; it += 3) {
            if (it + 2 >= a.nch) {
                comp(d0, k0, q0, w0);
                if (it + 1 < a.nch) comp(d1, k1, q1, w1);
                break;
            }
            locate(it + 2, d2, k2);
            load(d2, k2, q2, w2);
            comp(d0, k0, q0, w0);
            if (it + 3 >= a.nch) {
                comp(d1, k1, q1, w1);
                comp(d2, k2, q2, w2);
                break;
            }
            locate(it + 3, d0, k0);
            load(d0, k0, q0, w0);
            comp(d1, k1, q1, w1);
            if (it + 4 >= a.nch) {
                comp(d2, k2, q2, w2);
                comp(d0, k0, q0, w0);
                break;
            }
            locate(it + 4, d1, k1);
            load(d1, k1, q1, w1);
            comp(d2, k2, q2, w2);
        }
    }
    acc.x += alpha * lr.x;
    acc.y += alpha * lr.y;
    if (g > 0) red[g - 1][lane] = acc;
    __syncthreads();
    if (g == 0) {
        double2 r;
        r.x = acc.x + red[0][lane].x + red[1][lane].x + red[2][lane].x;
        r.y = acc.y + red[0][lane].y + red[1][lane].y + red[2][lane].y;
        double2 *yp = (double2 *)(y + i);
        if (beta != 0.0) {
            const double2 y0 = *yp;
            r.x += beta * y0.x;
            r.y += beta * y0.y;
        }
        *yp = r;
    }
}

static void mv2_tab(const dplrh_plan *p, int trans, Mv2Tab *t)
{
    memset(t, 0, sizeof(*t));
    t->L = p->L;
    t->N = p->N;
    t->trans = trans;
    int nb = 0, part = 0, fin = 0;
    for (int d = 0; d < p->L; ++d) {
        const int seg = p->N >> (d + 1), nseg = 2 << d, R = p->R[d];
        const int tps = seg <= 128 ? 1 : seg / 128;
        const int nkg = (R + KG - 1) / KG;
        t->R[d] = R;
        t->seg[d] = seg;
        t->TW[d] = 128;
        t->tps[d] = tps;
        t->nkg[d] = nkg;
        t->task_off[d] = nb;
        nb += nkg;
        t->part_off[d] = part;
        if (tps > 1) part += nseg * tps * R;
        t->fin_off[d] = fin;
        if (tps > 1) fin += nseg * R;
        t->w_off[d] = (int)p->soff[d];
        const double *Ud = p->U + p->uoff[d], *Vd = p->V + p->uoff[d];
        t->P[d] = trans ? Ud : Vd;
        t->Q[d] = trans ? Vd : Ud;
        t->S[d] = p->S + p->soff[d];
    }
    t->nb = nb;
    t->nch = 0;
    for (int d = 0; d < p->L; ++d) t->nch += (p->R[d] + 31) / 32;
    t->task_off[p->L] = nb;
    t->part_off[p->L] = part;
    t->fin_off[p->L] = fin;
    t->w_off[p->L] = (int)p->soff[p->L];
}

// Fast-path matvec for nc columns; the partial buffer p->part holds
// part_per_col doubles per column (sized by plr_workspace for this layout).
static int mv2_launch_ws(dplrh_plan *p, int trans, int nc, double alpha, double beta,
                         const double *x, int ldx, double *y, int ldy, double *part,
                         size_t pstride, double *w, size_t wstride, hipStream_t st)
{
    Mv2Tab t;
    mv2_tab(p, trans, &t);
    if ((size_t)t.part_off[t.L] > pstride) {
        plr_set_error("matvec workspace too small");
        return -9;
    }
    if (t.nb > 0) {
        hipLaunchKernelGGL(k_mv_reduce2, dim3(p->N / 128, nc), dim3(256), 0, st, t, x, ldx, part,
                           pstride, w, wstride);
        PLR_HIP(hipGetLastError());
    }
    if (t.fin_off[t.L] > 0) {
        hipLaunchKernelGGL(k_mv_finish, dim3((t.fin_off[t.L] + 31) / 32, nc), dim3(256), 0, st, t,
                           part, pstride, w, wstride);
        PLR_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mv_expand2, dim3(p->N / 128, nc), dim3(256), 0, st, t, p->b, p->D, x, ldx,
                       alpha, beta, trans, w, wstride, y, ldy);
    PLR_HIP(hipGetLastError());
    return 0;
}

static int mv2_launch(dplrh_plan *p, int trans, int nc, double alpha, double beta,
                      const double *x, int ldx, double *y, int ldy, hipStream_t st)
{
    return mv2_launch_ws(p, trans, nc, alpha, beta, x, ldx, y, ldy, p->part, p->part_per_col, p->w,
                         p->w_per_col, st);
}

int plr_launch_reduce(const LevelTab &t, const double *x, int ldx, double alpha, int trans,
                      int raw, int ncols, double *part, size_t pstride, int *cnt, size_t cstride,
                      double *w, size_t wstride, hipStream_t st)
{
    const int ntask = t.task_off[t.L];
    if (ntask <= 0 || ncols <= 0) return 0;
    dim3 grid((ntask + 3) / 4, ncols);
    hipLaunchKernelGGL(k_mv_reduce, grid, dim3(256), 0, st, t, x, ldx, alpha, trans, raw, part,
                       pstride, cnt, cstride, w, wstride);
    PLR_HIP(hipGetLastError());
    return 0;
}

int plr_launch_expand(const LevelTab &t, const double *D, const double *x, int ldx, double alpha,
                      double beta, int trans, const double *w, size_t wstride, double *y, int ldy,
                      int ncols, hipStream_t st)
{
    if (t.N <= 0 || ncols <= 0) return 0;
    dim3 grid((t.N + 63) / 64, ncols);
    hipLaunchKernelGGL(k_mv_expand, grid, dim3(256), 0, st, t, D, x, ldx, alpha, beta, trans, w,
                       wstride, y, ldy);
    PLR_HIP(hipGetLastError());
    return 0;
}

// y := alpha*op(A)*x + beta*y for ncols columns (alpha != 0), in chunks of
// at most kColChunk columns so the workspace and gridDim.y stay bounded.
static constexpr int kColChunk = 64;

int plr_mv_launch(dplrh_plan *p, int trans, int ncols, double alpha, double beta,
                  const double *dx, int ldx, double *dy, int ldy, hipStream_t st)
{
    const int chunk = ncols < kColChunk ? ncols : kColChunk;
    int rc = plr_workspace(p, chunk);
    if (rc) return rc;
    const bool fast = plr_mm_ok(p) && !p->force_valu && ldx % 2 == 0 && ldy % 2 == 0;
    if (fast) {
        for (int c0 = 0; c0 < ncols; c0 += chunk) {
            const int nc = ncols - c0 < chunk ? ncols - c0 : chunk;
            rc = mv2_launch(p, trans, nc, alpha, beta, dx + (size_t)c0 * ldx, ldx,
                            dy + (size_t)c0 * ldy, ldy, st);
            if (rc) return rc;
        }
        return 0;
    }
    LevelTab t;
    plr_level_tab(p, trans, &t);
    const int ntask = t.task_off[t.L];
    for (int c0 = 0; c0 < ncols; c0 += chunk) {
        const int nc = ncols - c0 < chunk ? ncols - c0 : chunk;
        const double *x = dx + (size_t)c0 * ldx;
        double *y = dy + (size_t)c0 * ldy;
        if (ntask > 0) {
            dim3 grid((ntask + 3) / 4, nc);
            hipLaunchKernelGGL(k_mv_reduce, grid, dim3(256), 0, st, t, x, ldx, alpha, trans, 0,
                               p->part, p->part_per_col, p->cnt, p->cnt_per_col, p->w,
                               p->w_per_col);
            PLR_HIP(hipGetLastError());
        }
        dim3 grid2((p->N + 63) / 64, nc);
        hipLaunchKernelGGL(k_mv_expand, grid2, dim3(256), 0, st, t, p->D, x, ldx, alpha, beta,
                           trans, p->w, p->w_per_col, y, ldy);
        PLR_HIP(hipGetLastError());
    }
    return 0;
}

static int gemv_body(dplrh_plan *p, int trans, double alpha, double beta, const double *dx,
                     double *dy, hipStream_t st)
{
    if (alpha == 0.0) return dplrh_scale2(p->N, 1, beta, dy, p->N, st);
    return plr_mv_launch(p, trans, 1, alpha, beta, dx, p->N, dy, p->N, st);
}

extern "C" int dplrh_gemv(dplrh_plan *p, int trans, double alpha, double beta,
                          const double *dx, double *dy, void *stream)
{
    if (p->N == 0) return 0;
    // Launched directly: three kernels finish sooner than a graph launch of
    // them (the graph's own start costs about as much as a kernel boundary).
    return gemv_body(p, trans, alpha, beta, dx, dy, (hipStream_t)stream);
}

static constexpr int kMmChunk = 128;

int plr_gemm_direct(dplrh_plan *p, int trans, int n, double alpha, double beta,
                    const double *dB, int ldb, double *dC, int ldc, hipStream_t st)
{
    if (alpha == 0.0) return dplrh_scale(p->N, n, beta, dC, ldc, st);
    if (n >= 2 && plr_mm_ok(p) && ldb % 2 == 0 && !p->force_valu) {
        const int chunk = n < kMmChunk ? n : kMmChunk;
        // whole 64-column groups: the multi-RHS product keeps its
        // coefficients rank-major over padded columns
        int rc = plr_workspace(p, (chunk + 63) & ~63);
        if (rc) return rc;
        for (int c0 = 0; c0 < n; c0 += chunk) {
            const int nc = n - c0 < chunk ? n - c0 : chunk;
            rc = plr_mm_apply(p, trans, nc, alpha, beta, dB + (size_t)c0 * ldb, ldb,
                              dC + (size_t)c0 * ldc, ldc, p->w, p->w_per_col, st);
            if (rc) return rc;
        }
        return 0;
    }
    return plr_mv_launch(p, trans, n, alpha, beta, dB, ldb, dC, ldc, st);
}

extern "C" int dplrh_gemm(dplrh_plan *p, int trans, int n, double alpha, double beta,
                          const double *dB, int ldb, double *dC, int ldc, void *stream)
{
    if (p->N == 0 || n == 0) return 0;
    // launched directly, as the matvec (three kernels per 128 columns)
    return plr_gemm_direct(p, trans, n, alpha, beta, dB, ldb, dC, ldc, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Matvec that allocates nothing.  The partial sums, coefficients and arrival
// counters of its kernels live in `scratch`, a static device array of
// DPLRX_SCRATCH_BYTES per GPU (each GPU the library runs on holds it until
// the process ends).  Its device address comes from hipGetSymbolAddress,
// once per device with that device current; host code only ever uses that
// address.  Every call on one GPU shares the array: calls that overlap in
// time on one GPU race.  Counters are cleared by a fill kernel on every call
// (the array keeps whatever the last call left).

static __device__ __align__(256) double scratch[DPLRX_SCRATCH_BYTES / sizeof(double)];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices];

static int scratch_on_device(double **out)
{
    int dev = 0;
    PLR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        plr_set_error("device %d is beyond the %d the scratch table holds", dev, kMaxDevices);
        return -5;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        PLR_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (double *)ptr;
    }
    *out = g_scratch[dev];
    return 0;
}

__global__ void k_zero_i32(int *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

extern "C" int dplrh_gemv_noalloc(dplrh_plan *p, int trans, double alpha, double beta,
                                  const double *dx, double *dy, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (p->N == 0) return 0;
    if (alpha == 0.0) return dplrh_scale2(p->N, 1, beta, dy, p->N, st);
    size_t part, w, cnt;
    plr_ws_per_col(p, &part, &w, &cnt);
    // 256-byte aligned regions: partials, coefficients, counters (ints)
    const size_t part_r = (part + 31) & ~(size_t)31, w_r = (w + 31) & ~(size_t)31;
    const size_t cnt_d = ((cnt + 1) / 2 + 31) & ~(size_t)31;
    const size_t need = part_r + w_r + cnt_d;
    if (need * sizeof(double) > DPLRX_SCRATCH_BYTES) {
        plr_set_error("matvec needs %zu bytes of scratch, more than the %zu the no-allocation "
                      "entry point holds", need * sizeof(double), (size_t)DPLRX_SCRATCH_BYTES);
        return -7;
    }
    double *base = nullptr;
    int rc = scratch_on_device(&base);
    if (rc) return rc;
    double *pp = base, *wp = base + part_r;
    int *cp = (int *)(wp + w_r);
    if (plr_mm_ok(p) && !p->force_valu)
        return mv2_launch_ws(p, trans, 1, alpha, beta, dx, p->N, dy, p->N, pp, part, wp, w, st);
    LevelTab t;
    plr_level_tab(p, trans, &t);
    const int ntask = t.task_off[t.L];
    if (ntask > 0) {
        if (cnt > 0) {
            hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                               cp, cnt);
            PLR_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_mv_reduce, dim3((ntask + 3) / 4, 1), dim3(256), 0, st, t, dx, p->N,
                           alpha, trans, 0, pp, part, cp, cnt, wp, w);
        PLR_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_mv_expand, dim3((p->N + 63) / 64, 1), dim3(256), 0, st, t, p->D, dx, p->N,
                       alpha, beta, trans, wp, w, dy, p->N);
    PLR_HIP(hipGetLastError());
    return 0;
}
