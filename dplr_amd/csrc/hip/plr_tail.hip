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
// When Rn > 0 the kernel also leaves T = V_{d0-1}^T z of its span (one
// segment of level d0 - 1) for the upper levels (k_fp_level).
#include <cstring>

#include "plr_common.h"
#include "plr_mm.h"

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
    int Rn;           // rank of level d0 - 1 when its T is wanted (multiple of 16), else 0
    double *Tout;     // T[c][s * Rn + k] of level d0 - 1, column stride tstride
    size_t tstride;
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

    // 4. T of level d0 - 1 over the span: tasks (16-rank tile, 32-row chunk),
    // chunk partials combined in a fixed order through the (free) T region
    if (a.Rn > 0) {
        const int Rn = a.Rn, nkt = Rn / 16, nch = span / 32, ntask = nkt * nch;
        const double *Vn = a.V[a.d0 - 1];
        double *red = T; // [task][lane][4]
        for (int task = wv; task < ntask; task += TL_NW) {
            const int kt = task / nch, ch = task - kt * nch;
            const double *pv = Vn + (size_t)(kt * 16 + m) * N + row0 + ch * 32 + g;
            double av[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) av[u] = pv[4 * u];
            tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = tl_mfma(av[u], z[m * zs + ch * 32 + 4 * u + g], acc);
            // lane holds T[kt*16 + g + 4v][column m]
#pragma unroll
            for (int v = 0; v < 4; ++v) red[(task * 64 + lane) * 4 + v] = acc[v];
        }
        __syncthreads();
        const int sg = row0 / span;
        for (int kt = wv; kt < nkt; kt += TL_NW) {
            double t[4] = {0.0, 0.0, 0.0, 0.0};
            for (int ch = 0; ch < nch; ++ch)
#pragma unroll
                for (int v = 0; v < 4; ++v) t[v] += red[((kt * nch + ch) * 64 + lane) * 4 + v];
            if (m < nc)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    a.Tout[(size_t)(c0 + m) * a.tstride + (size_t)sg * Rn + kt * 16 + g + 4 * v] = t[v];
        }
    }

    // 5. write the span back
    for (int e = threadIdx.x; e < TL_CW * span; e += 64 * TL_NW) {
        const int c = e / span, r = e - c * span;
        if (c < nc) X[(size_t)(c0 + c) * ldx + row0 + r] = z[c * zs + r];
    }
}

// ---------------------------------------------------------------------------
// The same bottom solve with its global loads issued a phase ahead, for
// ranks that are multiples of 16 and at most 32 (RT <= 32), with NW = span/32
// waves (two per 64-row chunk).  Work of a phase is spread over all waves:
//   leaves:  tasks (leaf, pair of 16-row tiles), the leaf's z read by every
//            task before any of them writes it back (a barrier between);
//   reduce:  tasks (64-row chunk of the span, rank tile): partial T per
//            chunk in the T region, [chunk][k][c], then summed per segment
//            (in chunk order) into the segment's first chunk; each lane
//            group takes rows 8h + 2g + {0,1} of a 16-row step, so one load
//            of the four groups covers 64 contiguous bytes of a V row;
//   cap:     tasks (node, 16-row tile of [c1; c2]);
//   expand:  tasks (16-row tile of the span), two per wave.
// The V rows of a level's reduction are loaded during the previous phase,
// inv(K) and sig during the reduction, Ut during the capacitance solve.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_solve_tail2(TailTab a, const double *__restrict__ B,
                                                     int ldb, double *__restrict__ X, int ldx,
                                                     int ncols)
{
    extern __shared__ double lds[];
    const int span = a.span, zs = span + 1, RT = a.RT, ts = TL_CW + 1;
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int blk = blockIdx.x, c0 = blockIdx.y * TL_CW;
    const int row0 = blk * span, N = a.N;
    const int nc = min(TL_CW, ncols - c0);
    const int nch = span / 64; // 64-row chunks (= leaves)
    double *z = lds;                  // [c][row], stride zs
    double *T = z + TL_CW * zs;       // [chunk][k][c], RT x ts per chunk
    double *W = T + nch * RT * ts;    // [segment][k][c]

    // leaf operands: task wv = (leaf wv/2, tiles 2(wv&1) + i); Dinv column-major 64 x 64
    const int lf = wv >> 1, qh = wv & 1;
    const bool lfw = lf < nch;
    double dv[2][16];
    {
        const double *Dl = a.Dinv + (size_t)(row0 / 64 + min(lf, nch - 1)) * 64 * 64;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int u = 0; u < 16; ++u) dv[i][u] = Dl[(size_t)(4 * u + g) * 64 + 16 * (2 * qh + i) + m];
    }
    // b of the span (columns c0.., zero padding columns), eight loads per
    // thread in flight at once
    for (int e0 = 0; e0 < TL_CW * span; e0 += 8 * 64 * NW) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = min(e0 + u * 64 * NW + (int)threadIdx.x, TL_CW * span - 1);
            const int c = e / span, r = e - c * span;
            v[u] = B[(size_t)(c0 + min(c, nc - 1)) * ldb + row0 + r];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * 64 * NW + (int)threadIdx.x;
            if (e < TL_CW * span) {
                const int c = e / span, r = e - c * span;
                z[c * zs + r] = c < nc ? v[u] : 0.0;
            }
        }
    }
    __syncthreads();
    // leaves: z|_leaf = Dinv_leaf b|_leaf
    {
        tl_d4 acc[2] = {tl_d4{0.0, 0.0, 0.0, 0.0}, tl_d4{0.0, 0.0, 0.0, 0.0}};
        if (lfw)
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const double bv = z[m * zs + lf * 64 + 4 * u + g];
#pragma unroll
                for (int i = 0; i < 2; ++i) acc[i] = tl_mfma(dv[i][u], bv, acc[i]);
            }
        __syncthreads(); // every read of the leaves is done
        if (lfw)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int v = 0; v < 4; ++v) z[m * zs + lf * 64 + 16 * (2 * qh + i) + g + 4 * v] = acc[i][v];
    }

    // reduction operands of level d: task wv = (chunk wv/2, rank tile wv&1);
    // va[q][h] = V[kt*16 + m][row0 + 64ch + 16q + 8h + 2g + {0,1}]
    const int rch = wv >> 1, rkt = wv & 1;
    double2 va[4][2];
    auto vload = [&](const double *Vd, int R) {
        const double *pv = Vd + (size_t)(min(rkt * 16, R - 16) + m) * N + row0 + 64 * min(rch, nch - 1) + 2 * g;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int h = 0; h < 2; ++h) va[q][h] = *(const double2 *)(pv + 16 * q + 8 * h);
    };
    int d = a.L - 1;
    while (d >= a.d0 && a.R[d] == 0) --d;
    if (d >= a.d0) vload(a.V[d], a.R[d]);
    __syncthreads();

    for (; d >= a.d0; --d) {
        const int R = a.R[d];
        if (R == 0) continue;
        const int seg = N >> (d + 1), nsg = span / seg, cps = seg / 64, nkt = R / 16, n2 = 2 * R;
        const int sg0 = row0 / seg; // global index of the span's first segment
        // 1. partial T per 64-row chunk
        if (rch < nch && rkt < nkt) {
            tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int row = 64 * rch + 16 * q + 8 * h + 2 * g + e;
                        acc = tl_mfma(e ? va[q][h].y : va[q][h].x, z[m * zs + row], acc);
                    }
            // lane holds T[rank rkt*16 + g + 4v][column m] of chunk rch
#pragma unroll
            for (int v = 0; v < 4; ++v) T[(rch * RT + rkt * 16 + g + 4 * v) * ts + m] = acc[v];
        }
        // operands of the capacitance solve: task wv = (node, 16-row tile it)
        const int ntl = n2 / 16, cnode = wv / ntl, cit = wv - cnode * ntl;
        const bool capw = cnode < nsg / 2;
        double kv[16], sv[4];
        {
            const int nodeg = (sg0 >> 1) + min(cnode, nsg / 2 - 1);
            const double *K = a.Kinv[d] + (size_t)nodeg * n2 * n2 + cit * 16 + m;
#pragma unroll
            for (int u = 0; u < 16; ++u) kv[u] = K[(size_t)min(4 * u + g, n2 - 1) * n2];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ii = cit * 16 + g + 4 * v;
                const int sgl = 2 * min(cnode, nsg / 2 - 1) + (ii < R ? 0 : 1);
                sv[v] = a.S[d][(size_t)(sg0 + sgl) * R + (ii < R ? ii : ii - R)];
            }
        }
        __syncthreads();
        // T of each segment: its chunks summed in order into its first chunk
        if (cps > 1) {
            for (int e = threadIdx.x; e < nsg * R * TL_CW; e += 64 * NW) {
                const int c = e % TL_CW, k = (e / TL_CW) % R, sgl = e / (TL_CW * R);
                double t = T[((sgl * cps) * RT + k) * ts + c];
                for (int cc = 1; cc < cps; ++cc) t += T[((sgl * cps + cc) * RT + k) * ts + c];
                T[((sgl * cps) * RT + k) * ts + c] = t;
            }
            __syncthreads();
        }
        // 2. c = inv(K) [T_bot; T_top]; W_top = -S_top .* c1, W_bot = -S_bot .* c2
        if (capw) {
            const int top = 2 * cnode, bot = top + 1; // local segments
            tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int j = 4 * u + g;
                if (4 * u < n2) {
                    const int sgl = j < R ? bot : top, kk = j < R ? j : j - R;
                    acc = tl_mfma(kv[u], T[((sgl * cps) * RT + kk) * ts + m], acc);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ii = cit * 16 + g + 4 * v;
                const int sgl = ii < R ? top : bot, kk = ii < R ? ii : ii - R;
                W[(sgl * RT + kk) * ts + m] = -sv[v] * acc[v];
            }
        }
        // operands of the expansion: tiles wv and wv + 16 of the span
        double uv[2][8];
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int tile = min(wv + NW * p, span / 16 - 1);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                uv[p][u] = a.Ut[d][(size_t)min(4 * u + g, R - 1) * N + row0 + 16 * tile + m];
        }
        __syncthreads();
        // 3. z|_s += Ut_d|_s W_s
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int tile = wv + NW * p;
            if (tile < span / 16) {
                const int rr = tile * 16, sl = rr / seg;
                tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (4 * u < R) acc = tl_mfma(uv[p][u], W[(sl * RT + 4 * u + g) * ts + m], acc);
                // lane holds (Ut W)[row rr + g + 4v][column m]
#pragma unroll
                for (int v = 0; v < 4; ++v) z[m * zs + rr + g + 4 * v] += acc[v];
            }
        }
        // the next level's reduction operands
        {
            int dn = d - 1;
            while (dn >= a.d0 && a.R[dn] == 0) --dn;
            if (dn >= a.d0) vload(a.V[dn], a.R[dn]);
        }
        __syncthreads();
    }

    // 4. T of level d0 - 1 over the span: tasks (16-rank tile, 32-row chunk),
    // chunk partials combined in a fixed order through the (free) T region
    if (a.Rn > 0) {
        const int Rn = a.Rn, nkt = Rn / 16, n32 = span / 32, ntask = nkt * n32;
        const double *Vn = a.V[a.d0 - 1];
        double *red = T; // [task][lane][4]
        for (int task = wv; task < ntask; task += NW) {
            const int kt = task / n32, ch = task - kt * n32;
            const double *pv = Vn + (size_t)(kt * 16 + m) * N + row0 + ch * 32 + g;
            double av[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) av[u] = pv[4 * u];
            tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int u = 0; u < 8; ++u) acc = tl_mfma(av[u], z[m * zs + ch * 32 + 4 * u + g], acc);
#pragma unroll
            for (int v = 0; v < 4; ++v) red[(task * 64 + lane) * 4 + v] = acc[v];
        }
        __syncthreads();
        const int sg = row0 / span;
        for (int kt = wv; kt < nkt; kt += NW) {
            double t[4] = {0.0, 0.0, 0.0, 0.0};
            for (int ch = 0; ch < n32; ++ch)
#pragma unroll
                for (int v = 0; v < 4; ++v) t[v] += red[((kt * n32 + ch) * 64 + lane) * 4 + v];
            if (m < nc)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    a.Tout[(size_t)(c0 + m) * a.tstride + (size_t)sg * Rn + kt * 16 + g + 4 * v] = t[v];
        }
    }

    // 5. write the span back
    for (int e = threadIdx.x; e < TL_CW * span; e += 64 * NW) {
        const int c = e / span, r = e - c * span;
        if (c < nc) X[(size_t)(c0 + c) * ldx + row0 + r] = z[c * zs + r];
    }
}

// ---------------------------------------------------------------------------
// One pass per level above the tail (d = d0-1 .. 0), FP64 MFMA throughout.
// A work-group owns 128 rows x 64 right-hand sides, four waves of 32 rows:
//   1. rhs = [T_bot; T_top] of its node into LDS
//   2. W = -S_s .* (the rows of inv(K) for its segment s) rhs   (tasks (kt, ct))
//   3. z^T += W^T Ut_d^T: a lane holds z[column g + 4v][row m] of each tile,
//      so z is read and written 16 consecutive rows (128 bytes) per lane group
//   4. z^T into LDS, then the partial V_{d-1}^T z over the 128 rows: wave w
//      takes rank tile w/2 and two column tiles; in MFMA step t of a 64-row
//      block lane group g contributes row 16g + t, so each lane streams 16
//      consecutive rows of one V column.  Written in mm_reduce_body's layout
//      for k_mm_finish (tiles of 128 rows).
// The old z and the Ut rows are loaded at the start, the V rows while the
// expansion runs.  Ranks R and Rn are multiples of 16 and at most 32.
static constexpr int FL_RB = 128, FL_CW = 64, FL_ZS = FL_RB + 1;

struct FpLvlTab
{
    int N, R, Rn, seg, segn, tps, ncg;
    const double *Ut, *Vn, *S, *Kinv;
};

static size_t fl_lds(int R)
{
    const size_t cap = (size_t)3 * R * (FL_CW + 1), zt = (size_t)FL_CW * FL_ZS;
    return (cap > zt ? cap : zt) * sizeof(double);
}

__global__ __launch_bounds__(256, 2) void k_fp_level(FpLvlTab a, const double *__restrict__ Tin,
                                                     size_t tstride, double *__restrict__ Z, int ldz,
                                                     double *__restrict__ part, int ncols)
{
    extern __shared__ double lds[];
    const int R = a.R, n2 = 2 * R, rs = FL_CW + 1;
    double *rhs = lds;         // [n2][rs]
    double *W = lds + n2 * rs; // [R][rs]: W[k][column]
    double *zt = lds;          // after the expansion: [column][FL_ZS]
    const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wv = tid >> 6;
    const int r0 = blockIdx.x * FL_RB, c0 = blockIdx.y * FL_CW, N = a.N;
    const int rw = 32 * wv; // this wave's rows (local)
    const int nc = min(FL_CW, ncols - c0);
    const int s = r0 / a.seg, istop = !(s & 1), top = s & ~1, bot = top + 1;
    const int nk4 = R / 4, nktn = a.Rn / 16;

    // old z: zo[rt][ct][v] = z[r0 + rw + 16rt + m][c0 + 16ct + g + 4v]
    double zo[2][4][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                zo[rt][ct][v] = Z[(size_t)(c0 + min(16 * ct + g + 4 * v, nc - 1)) * ldz + r0 + rw + 16 * rt + m];
    // Ut of the expansion: ua[rt][u] = Ut[4u + g][r0 + rw + 16rt + m]
    double ua[2][8];
    if (R > 0) // a level of rank 0 has no Ut rows
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
            for (int u = 0; u < 8; ++u)
                ua[rt][u] = a.Ut[(size_t)min(4 * u + g, R - 1) * N + r0 + rw + 16 * rt + m];
    // 1. rhs = [T_bot; T_top] of the node
    for (int e0 = 0; e0 < n2 * FL_CW; e0 += 8 * 256) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = min(e0 + u * 256 + tid, n2 * FL_CW - 1);
            const int cl = e / n2, j = e - cl * n2, hi = j >= R;
            const int sg = hi ? top : bot, kk = hi ? j - R : j;
            v[u] = Tin[(size_t)(c0 + min(cl, nc - 1)) * tstride + (size_t)sg * R + kk] *
                   (double)(cl < nc);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * 256 + tid;
            if (e < n2 * FL_CW) {
                const int cl = e / n2, j = e - cl * n2;
                rhs[j * rs + cl] = v[u];
            }
        }
    }
    __syncthreads();
    // 2. W[k][c] = -S_s[k] * sum_j inv(K)[row(k)][j] rhs[j][c]; row(k) = k (top) or R + k
    {
        const double *K = a.Kinv + (size_t)(s >> 1) * n2 * n2 + (istop ? 0 : R);
        const int ntask = (R / 16) * 4;
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int task = wv + 4 * p;
            if (task < ntask) {
                const int kt = task >> 2, ct = task & 3;
                double kv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    kv[u] = K[(size_t)min(4 * u + g, n2 - 1) * n2 + kt * 16 + m];
                tl_d4 acc = tl_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (4 * u < n2) acc = tl_mfma(kv[u], rhs[(4 * u + g) * rs + ct * 16 + m], acc);
                // lane holds c[kt*16 + g + 4v][ct*16 + m]
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = kt * 16 + g + 4 * v;
                    W[k * rs + ct * 16 + m] = -a.S[(size_t)s * R + k] * acc[v];
                }
            }
        }
    }
    __syncthreads();
    // V rows of the reduction (level d-1), in flight during the expansion:
    // wave w takes rank tile kt = w / 2; va[blk][h] = rows 64blk + 16g + 2h, +1
    const int kt = wv >> 1;
    double2 va[2][8];
    if (kt < nktn) {
        const double *pv = a.Vn + (size_t)(kt * 16 + m) * N + r0 + 16 * g;
#pragma unroll
        for (int blk = 0; blk < 2; ++blk)
#pragma unroll
            for (int h = 0; h < 8; ++h) va[blk][h] = *(const double2 *)(pv + 64 * blk + 2 * h);
    }
    // 3. z^T tiles: zn = zo + W^T Ut^T
    tl_d4 za[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) za[rt][ct] = tl_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        if (u < nk4) {
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const double w = W[(4 * u + g) * rs + ct * 16 + m];
#pragma unroll
                for (int rt = 0; rt < 2; ++rt) za[rt][ct] = tl_mfma(w, ua[rt][u], za[rt][ct]);
            }
        }
    }
    if (nktn > 0) __syncthreads(); // W is read by every wave before zt overwrites it
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = 16 * ct + g + 4 * v, i = rw + 16 * rt + m;
                const double zn = zo[rt][ct][v] + za[rt][ct][v];
                if (c < nc) Z[(size_t)(c0 + c) * ldz + r0 + i] = zn;
                if (nktn > 0) zt[c * FL_ZS + i] = zn;
            }
    if (nktn == 0) return;
    __syncthreads();
    // 4. partial V_{d-1}^T z over the 128 rows: tasks (kt, ct) with ct = 2(w&1), +1
    if (kt < nktn) {
        const int ct0 = 2 * (wv & 1);
        tl_d4 ta[2] = {tl_d4{0.0, 0.0, 0.0, 0.0}, tl_d4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int blk = 0; blk < 2; ++blk)
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const double av = (t & 1) ? va[blk][t >> 1].y : va[blk][t >> 1].x;
                const int row = 64 * blk + 16 * g + t;
#pragma unroll
                for (int q = 0; q < 2; ++q)
                    ta[q] = tl_mfma(av, zt[((ct0 + q) * 16 + m) * FL_ZS + row], ta[q]);
            }
        // lane holds T[kt*16 + g + 4v][(ct0+q)*16 + m]
        const int sn = r0 / a.segn, tile = (r0 - sn * a.segn) / FL_RB, st = sn * a.tps + tile;
        double *pp = part + ((size_t)st * a.ncg + blockIdx.y) * MM_BLK; // one rank group (Rn <= 32)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                pp[((ct0 + q) * 16 + m) * MM_KG + kt * 16 + g + 4 * v] = ta[q][v];
    }
}

// solve_tail_t: the shallowest level d0 it solved down to, TL_REFUSED when
// the shape is not covered (nothing launched), or a HIP error code (< -1).
static constexpr int TL_REFUSED = -1;
static int solve_tail_t(dplrh_plan *p, const double *Bm, int ldb, double *Xm, int ldx, int ncols,
                        int want_t, hipStream_t st);

int plr_solve_tail(dplrh_plan *p, const double *Bm, int ldb, double *Xm, int ldx, int ncols,
                   hipStream_t st)
{
    return solve_tail_t(p, Bm, ldb, Xm, ldx, ncols, 0, st);
}

// The whole non-transposed FP64 solve for ncols columns: k_solve_tail (with T
// of the first level above it), then per level k_fp_level and the finishing
// sums.  Returns 1, launching nothing, when the shape is not covered (leaf
// 64, N a multiple of 256, ranks above the tail multiples of 16 and at most
// 32, the workspace T sized by the caller for ncols).
int plr_tail_d0(const dplrh_plan *p)
{
    int d0 = 0;
    while ((p->N >> d0) > TL_MAXSPAN) ++d0;
    return d0;
}

int plr_solve_fast(dplrh_plan *p, const double *Bm, int ldb, double *Xm, int ldx, int ncols,
                   hipStream_t st)
{
    const int N = p->N, L = p->L;
    if (L == 0 || ncols <= 0 || p->b != 64 || N % 256 != 0) return 1;
    int d0 = 0;
    while ((N >> d0) > TL_MAXSPAN) ++d0;
    if (d0 == 0 || d0 > L - 1) return 1;
    for (int d = 0; d < d0; ++d)
        if (p->R[d] % 16 != 0 || p->R[d] > 32) return 1;
    int rc = plr_mm_part(p, plr_mm_part_blocks(p, 0, d0 - 1, ncols, FL_RB));
    if (rc) return rc;
    rc = solve_tail_t(p, Bm, ldb, Xm, ldx, ncols, 1, st);
    if (rc == TL_REFUSED) return 1; // the tail does not fit: nothing launched
    if (rc < 0) return rc;          // a HIP error, passed on
    // the reassociated upper levels when the factorization prepared them
    rc = plr_up_solve(p, d0, Xm, ldx, ncols, st);
    if (rc <= 0) return rc;
    static bool attr = false;
    if (!attr) {
        PLR_HIP(hipFuncSetAttribute((const void *)k_fp_level,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 80 * 1024));
        attr = true;
    }
    const int ncg = (ncols + FL_CW - 1) / FL_CW;
    for (int d = d0 - 1; d >= 0; --d) {
        FpLvlTab t;
        memset(&t, 0, sizeof(t));
        t.N = N;
        t.R = p->R[d];
        t.Rn = d > 0 ? p->R[d - 1] : 0;
        t.seg = N >> (d + 1);
        t.segn = N >> d;
        t.tps = t.segn / FL_RB;
        t.ncg = ncg;
        t.Ut = p->f.Ut + p->uoff[d];
        t.Vn = d > 0 ? p->V + p->uoff[d - 1] : nullptr;
        t.S = p->S + p->soff[d];
        t.Kinv = p->f.Kinv + p->f.koff[d];
        hipLaunchKernelGGL(k_fp_level, dim3(N / FL_RB, ncg), dim3(256), fl_lds(t.R), st, t, p->sw.T,
                           p->sw.t_per_col, Xm, ldx, p->tmp, ncols);
        PLR_HIP(hipGetLastError());
        if (t.Rn > 0 && (rc = plr_mm_finish_raw(p, d - 1, ncols, FL_RB, p->sw.T, p->sw.t_per_col, st)))
            return rc;
    }
    return 0;
}

static int solve_tail_t(dplrh_plan *p, const double *Bm, int ldb, double *Xm, int ldx, int ncols,
                        int want_t, hipStream_t st)
{
    const int N = p->N, L = p->L, b = p->b;
    if (L == 0 || ncols <= 0 || b != 64) return TL_REFUSED;
    // d0: the shallowest level whose span (two segments, N >> d0 rows) fits
    int d0 = 0;
    while ((N >> d0) > TL_MAXSPAN) ++d0;
    if (d0 > L - 1) return TL_REFUSED;
    const int span = N >> d0;
    int RT = 16;
    for (int d = d0 > 0 && want_t ? d0 - 1 : d0; d < L; ++d) {
        if (p->R[d] > TL_MAXR) return TL_REFUSED;
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
    if (want_t && d0 > 0) {
        t.Rn = p->R[d0 - 1];
        t.Tout = p->sw.T;
        t.tstride = p->sw.t_per_col;
    }
    const size_t tw = (size_t)(span / b) * RT * (TL_CW + 1);
    const size_t lds = ((size_t)TL_CW * (span + 1) + 2 * tw) * sizeof(double);
    static bool attr = false;
    if (!attr) {
        PLR_HIP(hipFuncSetAttribute((const void *)k_solve_tail,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    if (lds > 160 * 1024) return TL_REFUSED;
    dim3 grid(N / span, (ncols + TL_CW - 1) / TL_CW);
    // the second form: ranks of the tail levels multiples of 16 and at most 32
    bool second = RT <= 32 && (span == 512 || span == 256 || span == 128);
    for (int d = d0; d < L; ++d)
        if (p->R[d] % 16 != 0) second = false;
    if (second) {
        static bool attr2 = false;
        if (!attr2) {
            PLR_HIP(hipFuncSetAttribute((const void *)k_solve_tail2<16>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            PLR_HIP(hipFuncSetAttribute((const void *)k_solve_tail2<8>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            PLR_HIP(hipFuncSetAttribute((const void *)k_solve_tail2<4>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr2 = true;
        }
        if (span == 512)
            hipLaunchKernelGGL(k_solve_tail2<16>, grid, dim3(64 * 16), lds, st, t, Bm, ldb, Xm, ldx, ncols);
        else if (span == 256)
            hipLaunchKernelGGL(k_solve_tail2<8>, grid, dim3(64 * 8), lds, st, t, Bm, ldb, Xm, ldx, ncols);
        else
            hipLaunchKernelGGL(k_solve_tail2<4>, grid, dim3(64 * 4), lds, st, t, Bm, ldb, Xm, ldx, ncols);
    } else {
        hipLaunchKernelGGL(k_solve_tail, grid, dim3(64 * TL_NW), lds, st, t, Bm, ldb, Xm, ldx, ncols);
    }
    PLR_HIP(hipGetLastError());
    return d0;
}
