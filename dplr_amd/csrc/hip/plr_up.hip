// Upper levels of the FP64 solve, reassociated.
//
// Above the tail kernel (levels >= d0 and the leaves, solved span by span),
// the product form applies F_{d0-1}, ..., F_0 one level at a time:
//   T_d = V_d^T z,  W_d = cap_d(T_d),  z += Ut_d W_d,
// and every level reads and writes all of z.  Because Ut_e|_s is supported
// on segment s of level e, the T of a coarser level d < e only sees the
// update of level e through the small matrices
//   G_{d,e}[s] = V_d|_s^T Ut_e|_s          (R_d x R_e, one per segment s of level e)
// computed once at factorization:
//   T_d[s] = T0_d[s] + sum_{e > d} sum_{s' in s} G_{d,e}[s'] W_e[s'],   T0_d = V_d^T z_tail.
// The sums are carried up the tree as subtree totals
//   A_d^(e)[s] = G_{d,e}[s] W_e[s] + A_d^(e+1)[2s] + A_d^(e+1)[2s+1],
// so that T_d[s] = T0_d[s] + A_d^(d+1)[2s] + A_d^(d+1)[2s+1].  The solve is
// then one reduction pass over z (T0 of every upper level), one small
// kernel per level (T, the capacitance solve, the subtree totals), and one
// expansion pass (z += sum_d Ut_d W_d): z is read twice and written once
// instead of once per level.  The same quantities as the level-by-level
// order are computed, summed in another order.

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "plr_common.h"

typedef double up_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ up_d4 up_mfma(double a, double b, up_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static constexpr int UP_CW = 16;   // right-hand sides per work-group
static constexpr int UP_MAXR = 32; // ranks of the upper levels (multiples of 16)

struct UpTab
{
    int e, L, ncols;
    int R[PLR_MAXL];
    // T0 of level e: element (segment s, rank k, column c) at T + (s*R + k)*trs + c*tcs
    const double *T;
    size_t trs, tcs;
    const double *Ain;                  // A^(e+1) (null for the first level above the tail)
    double *Aout;                       // A^(e)
    size_t ain_off[PLR_MAXL], aout_off[PLR_MAXL]; // first row of level d's blocks
    double *W;                          // W_e rows soff[e] + s*R + k
    size_t w_row0, ws;                  // soff[e]; row stride of T0, W and A (columns)
    const double *Kinv, *S;             // level e
    const double *G[PLR_MAXL];          // G_{d,e}, d < e
};

// One work-group per (node of level e, 16 columns), eight waves.
//   1. rhs = [T_bot; T_top] with T = T0 + the children's subtree totals
//   2. c = inv(K) rhs (tiles of 16 rows); W_top = -S_top .* c1, W_bot = -S_bot .* c2
//   3. A_d^(e)[s] = G_{d,e}[s] W_e[s] + A_d^(e+1)[2s] + A_d^(e+1)[2s+1], d < e
// The kernel is latency-bound: every global operand of steps 2 and 3 (rows
// of inv(K), G, the children's totals) is loaded before step 1's barrier, so
// the work-group waits on memory about once.
static constexpr int UP_NW = 8; // waves per work-group
static constexpr int UP_QB = 4; // subtree-total tasks per wave prefetched at once

__device__ __forceinline__ void up_task(const UpTab &a, int task, int &d, int &hi, int &rt)
{
    int t = task;
    d = 0;
    while (t >= 2 * (a.R[d] / 16)) t -= 2 * (a.R[d++] / 16);
    const int nrt = a.R[d] / 16;
    hi = t / nrt;
    rt = t - hi * nrt;
}

__global__ __launch_bounds__(64 * UP_NW) void k_up_step(UpTab a)
{
    __shared__ double rhs[2 * UP_MAXR][UP_CW + 1];
    __shared__ double Wl[2][UP_MAXR][UP_CW + 1];
    const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int e = a.e, R = a.R[e], n2 = 2 * R;
    const int node = blockIdx.x, c0 = blockIdx.y * UP_CW;
    const int top = 2 * node, bot = top + 1;
    const size_t ws = a.ws;
    const int cm = min(c0 + m, a.ncols - 1); // this lane's column (clamped)

    // operands of step 2: rows of inv(K) (column-major n2 x n2) and sig for
    // tile wv of [c1; c2]
    const bool capw = wv < n2 / 16;
    double kv[2 * UP_MAXR / 4], sv[4];
    if (capw) {
        const double *K = a.Kinv + (size_t)node * n2 * n2 + wv * 16 + m;
#pragma unroll
        for (int u = 0; u < 2 * UP_MAXR / 4; ++u) kv[u] = K[(size_t)min(4 * u + g, n2 - 1) * n2];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = wv * 16 + g + 4 * v;
            const int hi = i >= R, sg = hi ? bot : top, k = hi ? i - R : i;
            sv[v] = a.S[(size_t)sg * R + k];
        }
    }
    // operands of step 3, first batch: tasks wv + 8q
    int ntask = 0;
    for (int d = 0; d < e; ++d) ntask += 2 * (a.R[d] / 16);
    double gv[UP_QB][UP_MAXR / 4], an[UP_QB][4];
    auto prefetch = [&](int q0) {
#pragma unroll
        for (int q = 0; q < UP_QB; ++q) {
            if (q0 + wv + UP_NW * q >= ntask) break;
            int d, hi, rt;
            up_task(a, q0 + wv + UP_NW * q, d, hi, rt);
            const int Rd = a.R[d], sg = hi ? bot : top;
            const double *Gs = a.G[d] + (size_t)sg * Rd * R + rt * 16 + m; // column-major Rd x R
#pragma unroll
            for (int u = 0; u < UP_MAXR / 4; ++u) gv[q][u] = Gs[(size_t)min(4 * u + g, R - 1) * Rd];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                an[q][v] = 0.0;
                if (a.Ain) {
                    const double *ap = a.Ain + (a.ain_off[d] + (size_t)(2 * sg) * Rd + rt * 16 + g + 4 * v) * ws + cm;
                    an[q][v] = ap[0] + ap[(size_t)Rd * ws];
                }
            }
        }
    };
    if (e > 0) prefetch(0);

    // 1. rhs[j][c]: j < R from the bottom segment, j >= R from the top one
    {
        double tv[2 * UP_MAXR * UP_CW / (64 * UP_NW)];
#pragma unroll
        for (int u = 0; u < 2 * UP_MAXR * UP_CW / (64 * UP_NW); ++u) {
            const int i = min(tid + 64 * UP_NW * u, n2 * UP_CW - 1);
            const int j = i / UP_CW, c = i - j * UP_CW;
            const int sg = j < R ? bot : top, k = j < R ? j : j - R;
            const int cc = min(c0 + c, a.ncols - 1);
            tv[u] = a.T[((size_t)sg * R + k) * a.trs + (size_t)cc * a.tcs];
            if (a.Ain) {
                const double *ap = a.Ain + (a.ain_off[e] + (size_t)(2 * sg) * R + k) * ws + cc;
                tv[u] += ap[0] + ap[(size_t)R * ws];
            }
        }
#pragma unroll
        for (int u = 0; u < 2 * UP_MAXR * UP_CW / (64 * UP_NW); ++u) {
            const int i = tid + 64 * UP_NW * u;
            if (i < n2 * UP_CW) {
                const int j = i / UP_CW, c = i - j * UP_CW;
                rhs[j][c] = c0 + c < a.ncols ? tv[u] : 0.0;
            }
        }
    }
    __syncthreads();

    // 2. the capacitance solve, 16 rows of [c1; c2] per wave
    if (capw) {
        up_d4 acc = up_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < 2 * UP_MAXR / 4; ++u)
            if (4 * u < n2) acc = up_mfma(kv[u], rhs[4 * u + g][m], acc);
        // lane holds c[wv*16 + g + 4v][column m]
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = wv * 16 + g + 4 * v;
            const int hi = i >= R, sg = hi ? bot : top, k = hi ? i - R : i;
            const double w = -sv[v] * acc[v];
            Wl[hi][k][m] = w;
            a.W[(a.w_row0 + (size_t)sg * R + k) * ws + c0 + m] = w;
        }
    }
    if (e == 0) return;
    __syncthreads();

    // 3. subtree totals for the coarser levels: tasks (d, segment, 16-row tile)
    for (int q0 = 0; q0 < ntask; q0 += UP_NW * UP_QB) {
        if (q0 > 0) prefetch(q0);
#pragma unroll
        for (int q = 0; q < UP_QB; ++q) {
            const int task = q0 + wv + UP_NW * q;
            if (task >= ntask) break;
            int d, hi, rt;
            up_task(a, task, d, hi, rt);
            const int Rd = a.R[d], sg = hi ? bot : top;
            up_d4 acc = up_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int u = 0; u < UP_MAXR / 4; ++u)
                if (4 * u < R) acc = up_mfma(gv[q][u], Wl[hi][4 * u + g][m], acc);
            // lane holds A[rt*16 + g + 4v][column m]
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = rt * 16 + g + 4 * v;
                a.Aout[(a.aout_off[d] + (size_t)sg * Rd + k) * ws + c0 + m] = acc[v] + an[q][v];
            }
        }
    }
}

// Shape of the reassociated path: leaf 64, N a multiple of 256, at least
// two levels above the tail, their ranks multiples of 16 and at most 32.
static int up_shape(const dplrh_plan *p)
{
    const int N = p->N, L = p->L;
    if (L < 3 || p->b != 64 || N % 256 != 0) return 0;
    const int d0 = plr_tail_d0(p);
    if (d0 < 2 || d0 > L - 1) return 0;
    for (int d = 0; d < d0; ++d)
        if (p->R[d] % 16 != 0 || p->R[d] > UP_MAXR || p->R[d] == 0) return 0;
    return d0;
}

// Rows (per column) of the subtree totals of level e: sum_{d<e} 2^{e+1} R_d.
static size_t up_arows(const dplrh_plan *p, int e, size_t *off)
{
    size_t r = 0;
    for (int d = 0; d < e; ++d) {
        if (off) off[d] = r;
        r += ((size_t)2 << e) * p->R[d];
    }
    return r;
}

// G_{d,e}[s][k + j R_d] = C[j * wstride + w_off + s * R_d + k]: the layout
// of plr_mm_cross's output turned into R_d x R_e column-major blocks.
__global__ void k_up_greshape(const double *__restrict__ C, size_t wstride, int w_off, int Rd,
                              int Re, int nseg, double *__restrict__ G)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t blk = (size_t)Rd * Re;
    if (e >= (size_t)nseg * blk) return;
    const size_t s = e / blk, r = e - s * blk;
    const int j = (int)(r / Rd), k = (int)(r - (size_t)j * Rd);
    G[e] = C[(size_t)j * wstride + w_off + s * Rd + k];
}

// All coarser levels of one level e at once, from the rank-major output of
// plr_mm_cross_rm: G_{d,e}[s][k + j R_d] = C[(w_off[d] + s R_d + k) * 64 + j].
struct UpGTab
{
    int e, Re, nseg;
    int Rd[PLR_MAXL], w_off[PLR_MAXL];
    size_t first[PLR_MAXL + 1]; // first element of level d in this launch
    double *G[PLR_MAXL];
};

__global__ void k_up_greshape_rm(const double *__restrict__ C, UpGTab t)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.first[t.e]) return;
    int d = 0;
    while (i >= t.first[d + 1]) ++d;
    const int Rd = t.Rd[d];
    const size_t r = i - t.first[d], blk = (size_t)Rd * t.Re;
    const size_t s = r / blk, q = r - s * blk;
    const int j = (int)(q / Rd), k = (int)(q - (size_t)j * Rd);
    t.G[d][r] = C[((size_t)t.w_off[d] + s * Rd + k) * 64 + j];
}

void plr_up_release(dplrh_plan *p) { (void)p; }

// Offsets of the G blocks of levels below d0 and their buffer (sizes
// depend on the shape only, so it is allocated once).
static int up_alloc_g(dplrh_plan *p, int d0)
{
    plr_fact &f = p->f;
    size_t tot = 0;
    for (int e = 1; e < d0; ++e)
        for (int d = 0; d < e; ++d) {
            f.goff[d][e] = tot;
            tot += ((size_t)2 << e) * p->R[d] * p->R[e];
        }
    if (!f.G) PLR_HIP(hipMalloc(&f.G, tot * sizeof(double)));
    return 0;
}

// G_{d,e} of every d < e from plr_mm_cross_rm's rank-major output C.
static int up_greshape(dplrh_plan *p, int e, const double *C, const int *w_off, hipStream_t st)
{
    plr_fact &f = p->f;
    const int Re = p->R[e], nseg = 2 << e;
    UpGTab g;
    memset(&g, 0, sizeof(g));
    g.e = e;
    g.Re = Re;
    g.nseg = nseg;
    size_t n = 0;
    for (int d = 0; d < e; ++d) {
        g.Rd[d] = p->R[d];
        g.w_off[d] = w_off[d];
        g.first[d] = n;
        g.G[d] = f.G + f.goff[d][e];
        n += (size_t)nseg * p->R[d] * Re;
    }
    g.first[e] = n;
    hipLaunchKernelGGL(k_up_greshape_rm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, C, g);
    PLR_HIP(hipGetLastError());
    return 0;
}

// After the factorization: the G blocks.  For each level e one reduction of
// Ut_e's columns against the V factors of every coarser level, cut into
// level e's segments (the FP64 MFMA reduction kernels with a table of those
// levels, tiles of 256 rows summed in a fixed order), then reshaped.  The
// shapes plr_up_begin does not take.
int plr_up_factor(dplrh_plan *p, hipStream_t st)
{
    plr_fact &f = p->f;
    const int d0 = up_shape(p);
    f.up_d0 = 0;
    if (!d0) return 0;
    const int N = p->N;
    int rc = up_alloc_g(p, d0);
    if (rc) return rc;
    double *C = f.work; // free once the factorization is done; it holds N * b doubles
    for (int e = 1; e < d0; ++e) {
        const int Re = p->R[e], nseg = 2 << e;
        int w_off[PLR_MAXL + 1];
        size_t ws = 0;
        rc = plr_mm_cross_rm(p, e, f.Ut + p->uoff[e], N, Re, C, (size_t)N * p->b, w_off, st);
        if (rc < 0) return rc;
        if (rc == 0) {
            if ((rc = up_greshape(p, e, C, w_off, st))) return rc;
            continue;
        }
        rc = plr_mm_cross(p, e, f.Ut + p->uoff[e], N, Re, C, w_off, &ws, st);
        if (rc) return rc;
        for (int d = 0; d < e; ++d) {
            const int Rd = p->R[d];
            const size_t n = (size_t)nseg * Rd * Re;
            hipLaunchKernelGGL(k_up_greshape, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, C,
                               ws, w_off[d], Rd, Re, nseg, f.G + f.goff[d][e]);
            PLR_HIP(hipGetLastError());
        }
    }
    f.up_d0 = d0;
    return 0;
}

// The G blocks beside the factorization's level loop.  Ut_e is final once
// the loop has finished level e + 1 (only finer levels update it), and the
// loop's later levels read it without writing it; so G of level e goes out
// on a second stream of default priority, after an event on the loop's
// stream, as level e's capacitance inverses start (a few work-groups each),
// with its own reduction output and partial sums.  plr_up_begin
// (before the loop) returns up_d0, or 0 when the shape takes plr_up_factor
// after the loop instead; plr_up_end makes the loop's stream wait for G.
int plr_up_begin(dplrh_plan *p)
{
    plr_fact &f = p->f;
    f.gd0 = 0;
    const int d0 = up_shape(p);
    if (!d0) return 0;
    size_t cmax = 0, pmax = 1;
    for (int e = 1; e < d0; ++e) {
        size_t c = 0, q = 0;
        if (plr_mm_cross_rm_size(p, e, &c, &q)) return 0;
        cmax = c > cmax ? c : cmax;
        pmax = q > pmax ? q : pmax;
    }
    int rc = up_alloc_g(p, d0);
    if (rc) return rc;
    if (!f.gC) PLR_HIP(hipMalloc(&f.gC, cmax * sizeof(double)));
    if (pmax > f.gpart_cap) {
        if (f.gpart) PLR_HIP(hipFree(f.gpart));
        f.gpart = nullptr;
        f.gpart_cap = 0;
        PLR_HIP(hipMalloc(&f.gpart, pmax * sizeof(double)));
        f.gpart_cap = pmax;
    }
    if (!f.gst) {
        PLR_HIP(hipStreamCreateWithFlags(&f.gst, hipStreamNonBlocking));
        PLR_HIP(hipEventCreateWithFlags(&f.gev, hipEventDisableTiming));
        PLR_HIP(hipEventCreateWithFlags(&f.gdone, hipEventDisableTiming));
    }
    f.gd0 = d0;
    return d0;
}

int plr_up_level(dplrh_plan *p, int e, hipStream_t st)
{
    plr_fact &f = p->f;
    if (!f.gd0 || e < 1 || e >= f.gd0) return 0;
    size_t c = 0, q = 0;
    if (plr_mm_cross_rm_size(p, e, &c, &q)) return -3;
    PLR_HIP(hipEventRecord(f.gev, st));
    PLR_HIP(hipStreamWaitEvent(f.gst, f.gev, 0));
    int w_off[PLR_MAXL + 1];
    const int rc = plr_mm_cross_rm(p, e, f.Ut + p->uoff[e], p->N, p->R[e], f.gC, c, w_off, f.gst,
                                   f.gpart, f.gpart_cap);
    if (rc) return rc < 0 ? rc : -3;
    return up_greshape(p, e, f.gC, w_off, f.gst);
}

int plr_up_end(dplrh_plan *p, hipStream_t st)
{
    plr_fact &f = p->f;
    if (!f.gd0) return 0;
    PLR_HIP(hipEventRecord(f.gdone, f.gst));
    PLR_HIP(hipStreamWaitEvent(st, f.gdone, 0));
    return 0;
}

static int up_ws(dplrh_plan *p, int d0, int ncols)
{
    plr_solve_ws &w = p->sw;
    const int cols = ((ncols + 63) / 64) * 64;
    if (cols <= w.ucols) return 0;
    ++p->ws_gen;
    void *old[] = {w.uT0, w.uW, w.uA[0], w.uA[1]};
    for (void *b : old)
        if (b) (void)hipFree(b);
    w.uT0 = w.uW = w.uA[0] = w.uA[1] = nullptr;
    w.ucols = 0;
    size_t arows = 0;
    for (int e = 1; e < d0; ++e) {
        const size_t r = up_arows(p, e, nullptr);
        arows = r > arows ? r : arows;
    }
    const size_t trows = p->soff[d0];
    PLR_HIP(hipMalloc(&w.uT0, trows * cols * sizeof(double)));
    PLR_HIP(hipMalloc(&w.uW, trows * cols * sizeof(double)));
    PLR_HIP(hipMalloc(&w.uA[0], (arows ? arows : 1) * cols * sizeof(double)));
    PLR_HIP(hipMalloc(&w.uA[1], (arows ? arows : 1) * cols * sizeof(double)));
    w.ucols = cols;
    return 0;
}

// The level steps e = d0-1 .. 0 (W of every level into p->sw.uW), with T0
// of level d0-1 in p->sw.T (column-major, t_per_col) when from_tail, else
// in p->sw.uT0 like the coarser levels.
static void up_steps(dplrh_plan *p, int d0, int ncols, int from_tail, hipStream_t st)
{
    plr_solve_ws &w = p->sw;
    const size_t ws = w.ucols;
    for (int e = d0 - 1; e >= 0; --e) {
        UpTab t;
        memset(&t, 0, sizeof(t));
        t.e = e;
        t.L = p->L;
        t.ncols = ncols;
        for (int d = 0; d < p->L; ++d) t.R[d] = p->R[d];
        if (e == d0 - 1 && from_tail) { // from the tail kernel: T[c * tstride + s*R + k]
            t.T = w.T;
            t.trs = 1;
            t.tcs = w.t_per_col;
        } else {
            t.T = w.uT0 + p->soff[e] * ws;
            t.trs = ws;
            t.tcs = 1;
        }
        if (e < d0 - 1) {
            t.Ain = w.uA[(e + 1) & 1];
            up_arows(p, e + 1, t.ain_off);
        }
        t.Aout = w.uA[e & 1];
        up_arows(p, e, t.aout_off);
        t.W = w.uW;
        t.w_row0 = p->soff[e];
        t.ws = ws;
        t.Kinv = p->f.Kinv + p->f.koff[e];
        t.S = p->S + p->soff[e];
        for (int d = 0; d < e; ++d) t.G[d] = p->f.G + p->f.goff[d][e];
        hipLaunchKernelGGL(k_up_step, dim3(1u << e, (ncols + UP_CW - 1) / UP_CW), dim3(64 * UP_NW),
                           0, st, t);
    }
}

// Levels d0-1 .. 0 of the solve over X (ldx), where the tail kernel left
// z_tail in X and T0 of level d0-1 in p->sw.T.  Returns 1, launching
// nothing, when the plan or X does not suit this path.
int plr_up_solve(dplrh_plan *p, int d0, double *X, int ldx, int ncols, hipStream_t st)
{
    if (!p->f.up_d0 || p->f.up_d0 != d0 || ldx % 2 != 0 || ((uintptr_t)X & 15) != 0) return 1;
    int rc = up_ws(p, d0, ncols);
    if (rc) return rc;
    plr_solve_ws &w = p->sw;
    const size_t ws = w.ucols;
    // T0 of levels 0 .. d0-2
    if ((rc = plr_mm_reduce_raw(p, d0 - 1, X, ldx, ncols, w.uT0, ws, st))) return rc;
    up_steps(p, d0, ncols, 1, st);
    PLR_HIP(hipGetLastError());
    return plr_mm_expand_w(p, d0, p->f.Ut, w.uW, ws, X, ldx, ncols, st);
}

// Whether the factorization prepared the level steps for levels below d0.
int plr_up_ready(const dplrh_plan *p, int d0) { return p->f.up_d0 >= d0 && d0 >= 2; }

// Workspace of the level steps for ncols columns (T0 and W: rows soff[e] +
// s*R + k of p->sw.uT0 and p->sw.uW, column stride p->sw.ucols).
int plr_up_reserve(dplrh_plan *p, int d0, int ncols) { return up_ws(p, d0, ncols); }

// The level steps alone, from T0 of every level 0 .. d0-1 in p->sw.uT0: W
// of every level in p->sw.uW.  For a solve that makes its own reductions and
// expansion (the mixed-precision one).
int plr_up_steps(dplrh_plan *p, int d0, int ncols, hipStream_t st)
{
    if (!plr_up_ready(p, d0)) return 1;
    up_steps(p, d0, ncols, 0, st);
    PLR_HIP(hipGetLastError());
    return 0;
}
