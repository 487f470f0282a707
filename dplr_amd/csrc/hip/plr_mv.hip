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
// the factors once).  Leaf orders 64 * 2^k take the two-launch form below
// (k_mv_reduce3 + k_mv_expand4/k_mv_expand3: no arrival counters, no fences).
#include <cstdint>
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

typedef double d2v __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// Matvec in two launches for leaf orders 64 * 2^k (every segment is 64 rows
// or a multiple of 128).  The U (and V) factors of all levels are one
// column-major N x sum(R_d) array, so a rank column of any level is one
// pointer step from the next.
//
// k_mv_reduce3: a work-group of eight waves owns MK rank columns of one level
// over 1024 rows (wave w: rows 128w..128w+127; lane l: rows 2l and 2l+1, one
// 16-byte load per column, all MK loads issued at once).  A wave sums its
// rows through an LDS transpose; segments of 64 and 128 rows are finished by
// the wave, 256, 512 and 1024 by two, four or eight waves in a fixed order,
// and longer segments leave one partial per 1024 rows.
// k_mv_expand3: a work-group of MW waves owns 128 rows.  It first stages
// alpha times the coefficients of its segments in LDS (summing the partials
// of segments longer than 1024 rows in a fixed order and scaling by sig),
// then each wave streams every MW-th column of [leaf | U of every level] in
// batches of UB 16-byte loads, and the waves' sums meet in LDS.

static constexpr int MV3_RB = 1024;   // rows of one reduction work-group
static constexpr int MV3_MAXSR = 1536; // most rank columns over all levels
static constexpr int MV3_MAXB = 1024;  // largest leaf order

struct Mv3Tab
{
    int L, N, nb, ng, SR, trans;
    int R[PLR_MAXL], seg[PLR_MAXL], kb_off[PLR_MAXL + 1], r_off[PLR_MAXL + 1];
    int w_off[PLR_MAXL + 1], p_off[PLR_MAXL + 1];
    const double *P, *Q, *S[PLR_MAXL];
};

// Sum of the ngs partials pp[j * R] of one coefficient (the 1024-row pieces
// of a long segment): even j into one sum and odd j into another, added at
// the end, with the loads of MV_PSUM partials issued before any is used
// (one L2 round trip per MV_PSUM partials, not per two; more would spill in
// the expansion, whose first two column batches are live here).
static constexpr int MV_PSUM = 8;
__device__ __forceinline__ double mv_part_sum(const double *__restrict__ pp, int ngs, int R)
{
    double v0 = 0.0, v1 = 0.0;
    for (int j0 = 0; j0 < ngs; j0 += MV_PSUM) {
        double t[MV_PSUM];
#pragma unroll
        for (int j = 0; j < MV_PSUM; ++j) t[j] = j0 + j < ngs ? pp[(size_t)(j0 + j) * R] : 0.0;
#pragma unroll
        for (int j = 0; j < MV_PSUM; j += 2) {
            v0 += t[j];
            v1 += t[j + 1];
        }
    }
    return v0 + v1;
}

template <int MK>
__global__ __launch_bounds__(512) void k_mv_reduce3(Mv3Tab a, const double *__restrict__ x,
                                                    int ldx, double *__restrict__ part,
                                                    size_t pstride, double *__restrict__ w,
                                                    size_t wstride)
{
    __shared__ double red[8][MK][65];
    __shared__ double sums[8][MK];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = blockIdx.y;
    const int grp = blockIdx.x % a.ng, bi = blockIdx.x / a.ng;
    int d = 0;
    while (bi >= a.kb_off[d + 1]) ++d;
    const int kl = (bi - a.kb_off[d]) * MK;
    const int N = a.N, R = a.R[d], kn = min(MK, R - kl), seg = a.seg[d];
    const int row0 = grp * MV3_RB + g * 128;
    const bool live = row0 < N;
    x += (size_t)col * ldx;
    part += (size_t)col * pstride;
    w += (size_t)col * wstride;
    if (live) {
        const int i = row0 + 2 * lane;
        const double *__restrict__ P = a.P + (size_t)(a.r_off[d] + kl) * N + i;
        d2v v[MK];
#pragma unroll
        for (int kk = 0; kk < MK; ++kk)
            v[kk] = __builtin_nontemporal_load((const d2v *)(P + (size_t)(kk < kn ? kk : kn - 1) * N));
        const d2v xv = *(const d2v *)(x + i);
#pragma unroll
        for (int kk = 0; kk < MK; ++kk) red[g][kk][lane] = v[kk].x * xv.x + v[kk].y * xv.y;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // lane (kk, pt) sums the MK lanes pt*MK .. pt*MK+MK-1 (2*MK rows)
    const int kk = lane % MK, pt = lane / MK;
    double sq = 0.0;
#pragma unroll
    for (int j = 0; j < MK; ++j) sq += red[g][kk][pt * MK + j];
#pragma unroll
    for (int off = MK; off < 32; off <<= 1) sq += __shfl_xor(sq, off, 64); // 64 rows
    const bool ok = kk < kn;
    const double *__restrict__ S = a.S[d];
    const int trans = a.trans;
    if (seg == 64) {
        if (live && ok && (lane & 31) < MK) {
            const int s = row0 / 64 + (lane >> 5);
            w[a.w_off[d] + (size_t)(s ^ 1) * R + kl + kk] =
                sq * S[(size_t)(trans ? s : s ^ 1) * R + kl + kk];
        }
        return;
    }
    sq += __shfl_xor(sq, 32, 64); // 128 rows
    if (seg == 128) {
        if (live && ok && lane < MK) {
            const int s = row0 / 128;
            w[a.w_off[d] + (size_t)(s ^ 1) * R + kl + kk] =
                sq * S[(size_t)(trans ? s : s ^ 1) * R + kl + kk];
        }
        return;
    }
    if (lane < MK) sums[g][lane] = live ? sq : 0.0;
    __syncthreads();
    const int t = threadIdx.x, cps = seg / 128;
    if (cps <= 8) {
        // groups of cps waves, one segment each
        const int gi = t / MK, k2 = t - gi * MK;
        if (gi < 8 / cps) {
            double v = 0.0;
            for (int j = 0; j < cps; ++j) v += sums[gi * cps + j][k2];
            const int r0 = grp * MV3_RB + gi * cps * 128;
            if (r0 < N && k2 < kn) {
                const int s = r0 / seg;
                w[a.w_off[d] + (size_t)(s ^ 1) * R + kl + k2] =
                    v * S[(size_t)(trans ? s : s ^ 1) * R + kl + k2];
            }
        }
    } else if (t < kn) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) v += sums[j][t];
        const int r0 = grp * MV3_RB, s = r0 / seg, ngs = seg / MV3_RB;
        const int gis = (r0 - s * seg) / MV3_RB;
        part[a.p_off[d] + ((size_t)s * ngs + gis) * R + kl + t] = v;
    }
}

template <int MW, int UB>
__global__ __launch_bounds__(MW * 64) void k_mv_expand3(Mv3Tab a, int b, const double *__restrict__ D,
                                                        const double *__restrict__ x, int ldx,
                                                        double alpha, double beta,
                                                        const double *__restrict__ part,
                                                        size_t pstride,
                                                        const double *__restrict__ w,
                                                        size_t wstride, double *__restrict__ y,
                                                        int ldy)
{
    __shared__ double coef[2][MV3_MAXSR]; // [row half][rank column]
    __shared__ double xs[2 * 64 > MV3_MAXB ? 2 * 64 : MV3_MAXB];
    __shared__ double2 red[MW - 1][64];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = blockIdx.y;
    x += (size_t)col * ldx;
    y += (size_t)col * ldy;
    part += (size_t)col * pstride;
    w += (size_t)col * wstride;
    const int N = a.N, SR = a.SR, trans = a.trans;
    const int row0 = blockIdx.x * 128, i = row0 + 2 * lane;
    const int leaf0 = row0 / b, nx = b == 64 ? 128 : b;
    // column c < b: leaf column c; c >= b: rank column c - b of every level.
    // Wave g takes columns c0 + g + MW * (UB * it + u), batch it of UB.
    const int h = lane >> 5; // row half (64-row segments, 64-row leaves)
    const int il = i - leaf0 * b - h * (b == 64 ? 64 : 0);
    const double *__restrict__ Dl = D + (size_t)(leaf0 + (b == 64 ? h : 0)) * b * b;
    const double *__restrict__ xl = xs + (b == 64 ? h * 64 : 0);
    const int ncol = (D ? b : 0) + SR, c0 = D ? 0 : b, cend = c0 + ncol;
    const int nbat = (ncol - g + MW * UB - 1) / (MW * UB);
    auto load = [&](d2v *v, int it) {
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int c = min(c0 + g + MW * (UB * it + u), cend - 1);
            const bool lf = c < b;
            const double *src = lf ? (trans ? Dl + (size_t)il * b + c : Dl + (size_t)c * b + il)
                                   : a.Q + (size_t)(c - b) * N + i;
            if (lf && trans) {
                v[u].x = src[0];
                v[u].y = src[b];
            } else {
                v[u] = __builtin_nontemporal_load((const d2v *)src);
            }
        }
    };
    double2 acc = {0.0, 0.0};
    auto comp = [&](const d2v *v, int it) {
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int c = c0 + g + MW * (UB * it + u);
            const int cc = min(c, cend - 1);
            const double m = cc < b ? xl[cc] : coef[h][cc - b];
            const double mu = c < cend ? m : 0.0;
            acc.x += v[u].x * mu;
            acc.y += v[u].y * mu;
        }
    };
    // the first batch is in flight while the coefficients are staged
    d2v va[UB], vb[UB];
    if (nbat > 0) load(va, 0);
    // alpha * x of the block's leaves, and alpha * the coefficients
    for (int t = threadIdx.x; t < nx; t += MW * 64) xs[t] = alpha * x[(size_t)leaf0 * b + t];
    for (int q = threadIdx.x; q < SR; q += MW * 64) {
        int d = 0;
        while (q >= a.r_off[d + 1]) ++d;
        const int k = q - a.r_off[d], R = a.R[d], seg = a.seg[d];
        if (seg == 64) {
            const int s = row0 / 64;
            coef[0][q] = alpha * w[a.w_off[d] + (size_t)s * R + k];
            coef[1][q] = alpha * w[a.w_off[d] + (size_t)(s + 1) * R + k];
        } else {
            const int s = row0 / seg;
            double v;
            if (seg > MV3_RB) {
                const int sp = s ^ 1, ngs = seg / MV3_RB;
                v = mv_part_sum(part + a.p_off[d] + (size_t)sp * ngs * R + k, ngs, R) *
                    a.S[d][(size_t)(trans ? sp : s) * R + k];
            } else {
                v = w[a.w_off[d] + (size_t)s * R + k];
            }
            coef[0][q] = coef[1][q] = alpha * v;
        }
    }
    __syncthreads();
    // two buffers in rotation: batch it + 1 is issued before batch it is
    // used, and every branch either leaves the loop or holds no load
    if (nbat > 0) {
        for (int it = 0;; it += 2) {
            if (it + 1 >= nbat) {
                comp(va, it);
                break;
            }
            load(vb, it + 1);
            comp(va, it);
            if (it + 2 >= nbat) {
                comp(vb, it + 1);
                break;
            }
            load(va, it + 2);
            comp(vb, it + 1);
        }
    }
    if (g > 0) red[g - 1][lane] = acc;
    __syncthreads();
    if (g == 0) {
        double2 r = acc;
#pragma unroll
        for (int j = 0; j < MW - 1; ++j) {
            r.x += red[j][lane].x;
            r.y += red[j][lane].y;
        }
        double2 *yp = (double2 *)(y + i);
        if (beta != 0.0) {
            const double2 y0 = *yp;
            r.x += beta * y0.x;
            r.y += beta * y0.y;
        }
        *yp = r;
    }
}

// The expansion over 256 rows per work-group (N a multiple of 256): lane l
// holds rows 2l, 2l+1 and 128 + 2l, 128 + 2l + 1 (two 16-byte loads per
// column), so every column is read as one 2 KB stretch per work-group.
template <int MW, int UB>
__global__ __launch_bounds__(MW * 64) void k_mv_expand4(Mv3Tab a, int b, const double *__restrict__ D,
                                                        const double *__restrict__ x, int ldx,
                                                        double alpha, double beta,
                                                        const double *__restrict__ part,
                                                        size_t pstride,
                                                        const double *__restrict__ w,
                                                        size_t wstride, double *__restrict__ y,
                                                        int ldy)
{
    __shared__ double coef[4][MV3_MAXSR]; // [64-row quarter][rank column]
    __shared__ double xs[256 > MV3_MAXB ? 256 : MV3_MAXB];
    __shared__ double2 red[MW - 1][2][64];
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col = blockIdx.y;
    x += (size_t)col * ldx;
    y += (size_t)col * ldy;
    part += (size_t)col * pstride;
    w += (size_t)col * wstride;
    const int N = a.N, SR = a.SR, trans = a.trans;
    const int row0 = blockIdx.x * 256;
    const int i0 = row0 + 2 * lane, i1 = i0 + 128;
    const int h0 = lane >> 5, h1 = 2 + h0; // 64-row quarters of the two row pairs
    const int xoff = b <= 256 ? row0 : row0 / b * b, nx = b <= 256 ? 256 : b;
    const int lf0 = i0 / b, lf1 = i1 / b;
    const int il0 = i0 - lf0 * b, il1 = i1 - lf1 * b;
    const double *__restrict__ D0 = D + (size_t)lf0 * b * b;
    const double *__restrict__ D1 = D + (size_t)lf1 * b * b;
    const double *__restrict__ x0 = xs + (lf0 * b - xoff);
    const double *__restrict__ x1 = xs + (lf1 * b - xoff);
    // column c < b: leaf column c; c >= b: rank column c - b of every level.
    // Wave g takes columns c0 + g + MW * (UB * it + u), batch it of UB.
    const int ncol = (D ? b : 0) + SR, c0 = D ? 0 : b, cend = c0 + ncol;
    const int nbat = (ncol - g + MW * UB - 1) / (MW * UB);
    auto load = [&](d2v *v, int it) {
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int c = min(c0 + g + MW * (UB * it + u), cend - 1);
            const bool lf = c < b;
            if (lf && trans) {
                v[2 * u].x = D0[(size_t)il0 * b + c];
                v[2 * u].y = D0[(size_t)(il0 + 1) * b + c];
                v[2 * u + 1].x = D1[(size_t)il1 * b + c];
                v[2 * u + 1].y = D1[(size_t)(il1 + 1) * b + c];
            } else {
                const double *s0 = lf ? D0 + (size_t)c * b + il0 : a.Q + (size_t)(c - b) * N + i0;
                const double *s1 = lf ? D1 + (size_t)c * b + il1 : a.Q + (size_t)(c - b) * N + i1;
                v[2 * u] = __builtin_nontemporal_load((const d2v *)s0);
                v[2 * u + 1] = __builtin_nontemporal_load((const d2v *)s1);
            }
        }
    };
    double2 acc0 = {0.0, 0.0}, acc1 = {0.0, 0.0};
    auto comp = [&](const d2v *v, int it) {
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int c = c0 + g + MW * (UB * it + u);
            const int cc = min(c, cend - 1);
            const bool lf = cc < b;
            const double m0 = lf ? x0[cc] : coef[h0][cc - b];
            const double m1 = lf ? x1[cc] : coef[h1][cc - b];
            const bool in = c < cend;
            const double mu0 = in ? m0 : 0.0, mu1 = in ? m1 : 0.0;
            acc0.x += v[2 * u].x * mu0;
            acc0.y += v[2 * u].y * mu0;
            acc1.x += v[2 * u + 1].x * mu1;
            acc1.y += v[2 * u + 1].y * mu1;
        }
    };
    // the first two batches are in flight while the coefficients are staged
    // (a single-batch wave loads its batch twice)
    d2v va[2 * UB], vb[2 * UB];
    if (nbat > 0) {
        load(va, 0);
        load(vb, min(1, nbat - 1));
    }
    // alpha * x of the block's leaves, and alpha * the coefficients of the
    // segments holding each 64-row quarter
    for (int t = threadIdx.x; t < nx; t += MW * 64) xs[t] = alpha * x[(size_t)xoff + t];
    for (int q = threadIdx.x; q < SR; q += MW * 64) {
        int d = 0;
        while (q >= a.r_off[d + 1]) ++d;
        const int k = q - a.r_off[d], R = a.R[d], seg = a.seg[d];
        const int nsg = seg >= 256 ? 1 : 256 / seg, per = 4 / nsg;
        for (int j = 0; j < nsg; ++j) {
            const int s = row0 / seg + j;
            double v;
            if (seg > MV3_RB) {
                const int sp = s ^ 1, ngs = seg / MV3_RB;
                v = mv_part_sum(part + a.p_off[d] + (size_t)sp * ngs * R + k, ngs, R) *
                    a.S[d][(size_t)(trans ? sp : s) * R + k];
            } else {
                v = w[a.w_off[d] + (size_t)s * R + k];
            }
            for (int hh = 0; hh < per; ++hh) coef[j * per + hh][q] = alpha * v;
        }
    }
    __syncthreads();
    if (nbat > 0) {
        // batches it and it + 1 are in flight on entry; each buffer is
        // refilled right after it is used, and every branch either leaves
        // the loop or holds no load
        for (int it = 0;; it += 2) {
            if (it + 2 >= nbat) {
                comp(va, it);
                if (it + 1 < nbat) comp(vb, it + 1);
                break;
            }
            comp(va, it);
            load(va, it + 2);
            if (it + 3 >= nbat) {
                comp(vb, it + 1);
                comp(va, it + 2);
                break;
            }
            comp(vb, it + 1);
            load(vb, it + 3);
        }
    }
    if (g > 0) {
        red[g - 1][0][lane] = acc0;
        red[g - 1][1][lane] = acc1;
    }
    __syncthreads();
    if (g == 0) {
        double2 r0 = acc0, r1 = acc1;
#pragma unroll
        for (int j = 0; j < MW - 1; ++j) {
            r0.x += red[j][0][lane].x;
            r0.y += red[j][0][lane].y;
            r1.x += red[j][1][lane].x;
            r1.y += red[j][1][lane].y;
        }
        double2 *y0p = (double2 *)(y + i0), *y1p = (double2 *)(y + i1);
        if (beta != 0.0) {
            const double2 ya = *y0p, yb = *y1p;
            r0.x += beta * ya.x;
            r0.y += beta * ya.y;
            r1.x += beta * yb.x;
            r1.y += beta * yb.y;
        }
        *y0p = r0;
        *y1p = r1;
    }
}

// Shapes the two-launch matvec covers: leaf order 64 * 2^k up to MV3_MAXB,
// at least one level, and at most MV3_MAXSR rank columns over all levels.
static bool mv3_ok(const dplrh_plan *p)
{
    if (p->L < 1 || p->b < 64 || p->b > MV3_MAXB || (p->b & (p->b - 1)) != 0) return false;
    return p->uoff[p->L] / (size_t)p->N <= (size_t)MV3_MAXSR;
}

static void mv3_tab(const dplrh_plan *p, int trans, int mk, Mv3Tab *t)
{
    memset(t, 0, sizeof(*t));
    t->L = p->L;
    t->N = p->N;
    t->trans = trans;
    t->ng = (p->N + MV3_RB - 1) / MV3_RB;
    int nb = 0, part = 0, ro = 0;
    for (int d = 0; d < p->L; ++d) {
        const int seg = p->N >> (d + 1), R = p->R[d];
        t->R[d] = R;
        t->seg[d] = seg;
        t->kb_off[d] = nb;
        nb += (R + mk - 1) / mk;
        t->r_off[d] = ro;
        ro += R;
        t->w_off[d] = (int)p->soff[d];
        t->p_off[d] = part;
        if (seg > MV3_RB) part += (2 << d) * (seg / MV3_RB) * R;
        t->S[d] = p->S + p->soff[d];
    }
    t->nb = nb;
    t->SR = ro;
    t->kb_off[p->L] = nb;
    t->r_off[p->L] = ro;
    t->w_off[p->L] = (int)p->soff[p->L];
    t->p_off[p->L] = part;
    t->P = trans ? p->U : p->V;
    t->Q = trans ? p->V : p->U;
}

// Launch shapes (measured at N=65536, L=10, leaf 64, rank 32: reduction
// 28.5-29.3 us, expansion 34.7-35.5 us for 370.7 MB)
static constexpr int MV3_MK = 16, MV4_MW = 16, MV4_UB = 2, MV3_MW = 16, MV3_UB = 4;

static int mv3_launch_ws(dplrh_plan *p, int trans, int nc, double alpha, double beta,
                         const double *x, int ldx, double *y, int ldy, double *part,
                         size_t pstride, double *w, size_t wstride, hipStream_t st)
{
    Mv3Tab t;
    mv3_tab(p, trans, MV3_MK, &t);
    if ((size_t)t.p_off[t.L] > pstride) {
        plr_set_error("matvec workspace too small");
        return -9;
    }
    if (t.nb > 0) {
        hipLaunchKernelGGL(k_mv_reduce3<MV3_MK>, dim3(t.ng * t.nb, nc), dim3(512), 0, st, t, x,
                           ldx, part, pstride, w, wstride);
        PLR_HIP(hipGetLastError());
    }
    if (p->N % 256 == 0)
        hipLaunchKernelGGL((k_mv_expand4<MV4_MW, MV4_UB>), dim3(p->N / 256, nc),
                           dim3(MV4_MW * 64), 0, st, t, p->b, p->D, x, ldx, alpha, beta, part,
                           pstride, w, wstride, y, ldy);
    else
        hipLaunchKernelGGL((k_mv_expand3<MV3_MW, MV3_UB>), dim3(p->N / 128, nc),
                           dim3(MV3_MW * 64), 0, st, t, p->b, p->D, x, ldx, alpha, beta, part,
                           pstride, w, wstride, y, ldy);
    PLR_HIP(hipGetLastError());
    return 0;
}

static int mv3_launch(dplrh_plan *p, int trans, int nc, double alpha, double beta,
                      const double *x, int ldx, double *y, int ldy, hipStream_t st)
{
    return mv3_launch_ws(p, trans, nc, alpha, beta, x, ldx, y, ldy, p->part, p->part_per_col, p->w,
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
    // the fast kernels read x and y two rows at a time (16-byte loads)
    const bool fast = mv3_ok(p) && !p->force_valu && ldx % 2 == 0 && ldy % 2 == 0 &&
                      (((uintptr_t)dx | (uintptr_t)dy) & 15) == 0;
    if (fast) {
        for (int c0 = 0; c0 < ncols; c0 += chunk) {
            const int nc = ncols - c0 < chunk ? ncols - c0 : chunk;
            rc = mv3_launch(p, trans, nc, alpha, beta, dx + (size_t)c0 * ldx, ldx,
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
    const bool two = mv3_ok(p) && !p->force_valu && (((uintptr_t)dx | (uintptr_t)dy) & 15) == 0;
    if (two) {
        // the two-launch kernels use no counters and fewer partials
        Mv3Tab t;
        mv3_tab(p, trans, MV3_MK, &t);
        part = (size_t)t.p_off[t.L];
        cnt = 0;
    }
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
    if (two)
        return mv3_launch_ws(p, trans, 1, alpha, beta, dx, p->N, dy, p->N, pp, part, wp, w, st);
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
