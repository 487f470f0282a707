// The bottom of the factored FP64 solve in one kernel.
//
// z = inv(D_leaf) b, then F_{L-1}, ..., F_{d0} with d0 the shallowest level
// whose segments have at most 256 rows: these factors only mix rows inside a
// span of 2 * seg(d0) rows, so the solve is block-diagonal over spans.  One
// work-group per (span, 16 right-hand sides) keeps z of its span in LDS and
// runs every step on v_mfma_f64_16x16x4_f64:
//   leaves:   z|_leaf = Dinv_leaf b|_leaf                 (A = Dinv from HBM)
//   per level d = L-1 .. d0, per segment s of the span:
//     reduce: T_s = V_d|_s^T z|_s                          (A = V from HBM, B = z in LDS)
//     cap:    c = inv(K_node) [T_bot; T_top], W_top = -S_top .* c1, W_bot = -S_bot .* c2
//     expand: z|_s += Ut_d|_s W_s                         (A = Ut from HBM, B = W in LDS)
// The levels above d0 keep their per-level kernels.
#include <cstring>

#include "plr_common.h"

typedef double tl_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ tl_d4 tl_mfma(double a, double b, tl_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static constexpr int TL_CW = 16;     // right-hand sides per work-group
static constexpr int TL_NW = 16;     // waves per work-group (latency hiding: one group per CU)
static constexpr int TL_MAXSPAN = 512;
static constexpr int TL_MAXR = 64;   // ranks per level handled here

struct TailTab
{
    int N, b, L, d0, span, RT; // RT: rank stride of T and W (max tail rank, multiple of 16)
    int R[PLR_MAXL];
    const double *V[PLR_MAXL], *Ut[PLR_MAXL], *S[PLR_MAXL], *Kinv[PLR_MAXL];
    const double *Dinv;
};

// LDS: z[c][row] (row stride span + 1), T and W as [seg][k][c] (k stride 17)
__global__ __launch_bounds__(64 * TL_NW) void k_solve_tail(TailTab a, const double *__restrict__ B, int ldb,
                                                    double *__restrict__ X, int ldx, int ncols)
{
    extern __shared__ double lds[];
    const int span = a.span, zs = span + 1, RT = a.RT;
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4, wv = threadIdx.x >> 6;
    const int blk = blockIdx.x, c0 = blockIdx.y * TL_CW;
    const int row0 = blk * span, N = a.N, b = a.b;
    const int nc = min(TL_CW, ncols - c0);
    // z of the span, then the per-level T and W as [segment][k][column]
    // (k stride TL_CW + 1; at most span / b segments)
    double *z = lds;
    double *T = z + TL_CW * zs;
    double *W = T + (span / b) * RT * (TL_CW + 1);

    // 1. load b of the span (columns c0..c0+15), zero padding columns
    for (int e = threadIdx.x; e < TL_CW * span; e += 64 * TL_NW) {
        const int c = e / span, r = e - c * span;
        const double v = B[(size_t)(c0 + min(c, nc - 1)) * ldb + row0 + r];
        z[c * zs + r] = c < nc ? v : 0.0;
    }
    __syncthreads();

    // 2. leaves (b = 64): each wave owns whole leaves and computes all four
    // 16-row tiles of a leaf before writing it back in place
    const int nleaf = span / b;
    for (int l = wv; l < nleaf; l += TL_NW) {
        const double *Dl = a.Dinv + (size_t)(row0 / b + l) * b * b;
        const int lb = l * b;
        tl_d4 acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = tl_d4{0.0, 0.0, 0.0, 0.0};
        for (int j0 = 0; j0 < 64; j0 += 16) {
            double av[4][4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + 4 * u + g;
                bv[u] = z[m * zs + lb + j];
#pragma unroll
                for (int q = 0; q < 4; ++q) av[u][q] = Dl[(size_t)j * 64 + 16 * q + m];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = tl_mfma(av[u][q], bv[u], acc[q]);
        }
        // this wave's reads of the leaf are done (LDS operations of a wave
        // complete in order); lane holds z[16q + g + 4v][column m]
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) z[m * zs + lb + 16 * q + g + 4 * v] = acc[q][v];
    }
    __syncthreads();

    // 3. levels L-1 .. d0
    for (int d = a.L - 1; d >= a.d0; --d) {
        const int R = a.R[d];
        if (R == 0) continue;
        const int seg = N >> (d + 1), nsg = span / seg, nkt = (R + 15) / 16;
        const int sg0 = row0 / seg; // global index of the span's first segment
        const double *Vd = a.V[d], *Ud = a.Ut[d], *Sd = a.S[d];
        // reduce: task = (segment, rank tile); K = seg rows
        for (int task = wv; task < nsg * nkt; task += TL_NW) {
            const int sl = task / nkt, kt = task - sl * nkt;
            const int k = kt * 16 + m;
            const double pon = k < R ? 1.0 : 0.0;
            const double *pv = Vd + (size_t)min(k, R - 1) * N + row0 + sl * seg;
            tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
            for (int r = 0; r < seg; r += 64) {
                double av[16];
#pragma unroll
                for (int t = 0; t < 16; ++t) av[t] = pv[r + 16 * g + t];
#pragma unroll
                for (int t = 0; t < 16; ++t)
                    acc = tl_mfma(av[t] * pon, z[m * zs + sl * seg + r + 16 * g + t], acc);
            }
            // lane holds T[rank kt*16 + g + 4v][column m]
#pragma unroll
            for (int v = 0; v < 4; ++v) T[(sl * RT + kt * 16 + g + 4 * v) * (TL_CW + 1) + m] = acc[v];
        }
        __syncthreads();
        // cap: task = (node, 16-row tile of the 2R rows); K = 2R
        const int n2 = 2 * R, nrt = (n2 + 15) / 16, nodes = nsg / 2;
        for (int task = wv; task < nodes * nrt; task += TL_NW) {
            const int nl = task / nrt, it = task - nl * nrt;
            const int node = (sg0 >> 1) + nl, top = 2 * nl, bot = top + 1; // local segments
            const double *K = a.Kinv[d] + (size_t)node * n2 * n2;
            const int i = it * 16 + m, ia = min(i, n2 - 1);
            tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
            for (int j0 = 0; j0 < n2; j0 += 16) {
                double av[4], bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = min(j0 + 4 * u + g, n2 - 1);
                    const double on = (j0 + 4 * u + g < n2) ? 1.0 : 0.0;
                    av[u] = K[(size_t)j * n2 + ia] * on;
                    // rhs = [T_bot; T_top]
                    const int sgl = j < R ? bot : top, kk = j < R ? j : j - R;
                    bv[u] = T[(sgl * RT + kk) * (TL_CW + 1) + m];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) acc = tl_mfma(av[u], bv[u], acc);
            }
            // lane holds c[it*16 + g + 4v][column m]; W_top = -S_top .* c1, W_bot = -S_bot .* c2
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ii = it * 16 + g + 4 * v;
                if (ii >= n2) continue;
                const int sgl = ii < R ? top : bot, kk = ii < R ? ii : ii - R;
                const double sv = Sd[(size_t)(sg0 + sgl) * R + kk];
                W[(sgl * RT + kk) * (TL_CW + 1) + m] = -sv * acc[v];
            }
        }
        __syncthreads();
        // expand: task = 16-row tile of the span; K = R
        for (int task = wv; task < span / 16; task += TL_NW) {
            const int rr = task * 16, sl = rr / seg;
            tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < R; k0 += 16) {
                double av[4], bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + 4 * u + g, kc = min(k, R - 1);
                    const double on = k < R ? 1.0 : 0.0;
                    av[u] = Ud[(size_t)kc * N + row0 + rr + m] * on;
                    bv[u] = W[(sl * RT + kc) * (TL_CW + 1) + m];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) acc = tl_mfma(av[u], bv[u], acc);
            }
            // lane holds (Ut W)[row rr + g + 4v][column m]
#pragma unroll
            for (int v = 0; v < 4; ++v) z[m * zs + rr + g + 4 * v] += acc[v];
        }
        __syncthreads();
    }

    // 4. write the span back
    for (int e = threadIdx.x; e < TL_CW * span; e += 64 * TL_NW) {
        const int c = e / span, r = e - c * span;
        if (c < nc) X[(size_t)(c0 + c) * ldx + row0 + r] = z[c * zs + r];
    }
}

// The host side: handles the leaves and the levels d >= d0, returns d0 (the
// caller runs the levels above), or -1 when the shape does not fit.
int plr_solve_tail(dplrh_plan *p, const double *Bm, int ldb, double *Xm, int ldx, int ncols,
                   hipStream_t st)
{
    const int N = p->N, L = p->L, b = p->b;
    if (L == 0 || ncols <= 0 || b != 64) return -1;
    // d0: the shallowest level whose span (two segments, N >> d0 rows) fits
    int d0 = 0;
    while ((N >> d0) > TL_MAXSPAN) ++d0;
    if (d0 > L - 1) return -1;
    const int span = N >> d0;
    int RT = 16;
    for (int d = d0; d < L; ++d) {
        if (p->R[d] > TL_MAXR) return -1;
        RT = ((p->R[d] + 15) / 16) * 16 > RT ? ((p->R[d] + 15) / 16) * 16 : RT;
    }
    TailTab t;
    memset(&t, 0, sizeof(t));
    t.N = N;
    t.b = b;
    t.L = L;
    t.d0 = d0;
    t.span = span;
    t.RT = RT;
    for (int d = 0; d < L; ++d) {
        t.R[d] = p->R[d];
        t.V[d] = p->V + p->uoff[d];
        t.Ut[d] = p->f.Ut + p->uoff[d];
        t.S[d] = p->S + p->soff[d];
        t.Kinv[d] = p->f.Kinv + p->f.koff[d];
    }
    t.Dinv = p->f.Dinv;
    const size_t tw = (size_t)(span / b) * RT * (TL_CW + 1);
    const size_t lds = ((size_t)TL_CW * (span + 1) + 2 * tw) * sizeof(double);
    static bool attr = false;
    if (!attr) {
        PLR_HIP(hipFuncSetAttribute((const void *)k_solve_tail,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    if (lds > 160 * 1024) return -1;
    dim3 grid(N / span, (ncols + TL_CW - 1) / TL_CW);
    hipLaunchKernelGGL(k_solve_tail, grid, dim3(64 * TL_NW), lds, st, t, Bm, ldb, Xm, ldx, ncols);
    PLR_HIP(hipGetLastError());
    return d0;
}
