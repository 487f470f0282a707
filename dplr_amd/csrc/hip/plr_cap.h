// Capacitance step of the factored solve on FP64 MFMA, shared by the FP64
// and the mixed-precision paths.
//
// Per node n of level d (segments top = 2n, bot = 2n + 1, order n2 = 2R) and
// a chunk of up to 64 right-hand sides:
//   not transposed: rhs = [t_bot; t_top], c = inv(K) rhs,
//                   W_top = -S_top .* c[0:R], W_bot = -S_bot .* c[R:2R]
//   transposed:     rhs = [S_top .* t_top; S_bot .* t_bot], c = inv(K)^T rhs,
//                   W_bot = -c[0:R], W_top = -c[R:2R]
// inv(K) is column-major n2 x n2 per node.  One work-group per (node, column
// chunk); its four waves take 16-row tiles of c; the right-hand sides sit in
// LDS as rhs[j][column], inv(K) goes straight from HBM into the A operand in
// batches of 16 loads.  With LP the outputs are FP32 and each (segment,
// column) also gets the power-of-two scale of its largest magnitude.  Output
// element (row j, column c) goes to W[c * wstride + j * rs]: rs = 1 keeps
// the layout of T, wstride = 1 and rs = the column count is rank-major.
#pragma once

#include "plr_common.h"

typedef double plr_d4 __attribute__((ext_vector_type(4)));

template <typename KT, typename VT, bool LP, int NCT>
__device__ __forceinline__ void cap_body(int node, int cblk, const KT *__restrict__ Kinv, int R,
                                         const VT *__restrict__ S, const VT *__restrict__ T,
                                         size_t tstride, VT *__restrict__ W, size_t wstride,
                                         int trans, int ncols, float *__restrict__ wscale,
                                         size_t sstride, size_t rs = 1)
{
    // rhs[j][column], j < n2p (rows n2..n2p-1 zero), row stride cw + 1
    extern __shared__ double rhs[];
    __shared__ unsigned smax[64][2];
    constexpr int cw = 16 * NCT, cwp = cw + 1;
    const int c0 = cblk * cw, n2 = 2 * R, n2p = (n2 + 63) & ~63;
    const int nc = min(cw, ncols - c0);
    const int top = 2 * node, bot = top + 1;
    const KT *__restrict__ k = Kinv + (size_t)node * n2 * n2;
    // staging: consecutive threads take consecutive j of one column (contiguous
    // in T), eight independent loads per thread in flight
    const int ne = n2p * cw;
    for (int e0 = 0; e0 < ne; e0 += 8 * 256) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * 256 + (int)threadIdx.x;
            const int cl = e / n2p, j = e - cl * n2p;
            const int jc = min(j, n2 - 1), hi = jc >= R, kk = hi ? jc - R : jc;
            const int cc = c0 + min(cl, nc - 1);
            // not transposed: [t_bot; t_top]; transposed: [S.*t_top; S.*t_bot]
            const int sg = trans ? (hi ? bot : top) : (hi ? top : bot);
            // Loads are unconditional and padding is zeroed by a 0/1 factor: a
            // select that drops a load lets the compiler move the load into a
            // branch that waits for it alone.  Clamped indices read real data.
            const double tv = (double)T[(size_t)cc * tstride + (size_t)sg * R + kk];
            const double sv = (double)S[(size_t)sg * R + kk];
            v[u] = (trans ? sv * tv : tv) * (double)(cl < nc && j < n2);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * 256 + (int)threadIdx.x;
            if (e < ne) {
                const int cl = e / n2p, j = e - cl * n2p;
                rhs[(size_t)j * cwp + cl] = v[u];
            }
        }
    }
    if (LP)
        for (int e = threadIdx.x; e < 128; e += blockDim.x) smax[e >> 1][e & 1] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int ntile = (n2 + 15) / 16;
    const size_t ks_i = trans ? n2 : 1, ks_j = trans ? 1 : n2; // element (i, j) of op(inv(K))
    for (int it = wv; it < ntile; it += 4) {
        const int ia = min(it * 16 + m, n2 - 1);
        const bool arow = it * 16 + m < n2;
        plr_d4 acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[ct] = plr_d4{0.0, 0.0, 0.0, 0.0};
        for (int kb0 = 0; kb0 < n2p; kb0 += 64) {
            double av[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int j = kb0 + 4 * u + g, jc = min(j, n2 - 1);
                av[u] = (double)k[ia * ks_i + jc * ks_j] * (double)(arow && j < n2);
            }
            const double *bq = rhs + (size_t)(kb0 + g) * cwp + m;
#pragma unroll
            for (int u = 0; u < 16; ++u)
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct)
                    acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bq[4 * u * cwp + ct * 16],
                                                                   acc[ct], 0, 0, 0);
        }
        // lane holds c[it*16 + g + 4v][ct*16 + m]
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
            const int cl = ct * 16 + m;
            if (cl >= nc) continue;
            VT *w = W + (size_t)(c0 + cl) * wstride;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = it * 16 + g + 4 * v;
                if (i >= n2) continue;
                const double c = acc[ct][v];
                const int kk = i < R ? i : i - R;
                const int sg = (i < R) != (trans != 0) ? top : bot;
                const double sv = (double)S[(size_t)sg * R + kk];
                const VT out = (VT)(trans ? -c : -sv * c);
                w[((size_t)sg * R + kk) * rs] = out;
                if (LP) atomicMax(&smax[cl][i < R ? 0 : 1], __float_as_uint(fabsf((float)out)));
            }
        }
    }
    if (LP) {
        __syncthreads();
        for (int e = threadIdx.x; e < 2 * nc; e += blockDim.x) {
            const int cl = e >> 1, half = e & 1;
            const float mv = __uint_as_float(smax[cl][half]);
            // smallest power of two >= the largest magnitude (1 for an all-zero segment)
            const float sc = mv > 0.f ? exp2f(ceilf(log2f(mv))) : 1.0f;
            wscale[(size_t)(c0 + cl) * sstride + (half ? bot : top)] = sc;
        }
    }
}

template <typename KT, typename VT, bool LP, int NCT>
__global__ __launch_bounds__(256) void k_cap_mm(const KT *__restrict__ Kinv, int R,
                                                const VT *__restrict__ S,
                                                const VT *__restrict__ T, size_t tstride,
                                                VT *__restrict__ W, size_t wstride, int trans,
                                                int ncols, float *__restrict__ wscale,
                                                size_t sstride, size_t rs)
{
    cap_body<KT, VT, LP, NCT>(blockIdx.x, blockIdx.y, Kinv, R, S, T, tstride, W, wstride, trans,
                              ncols, wscale, sstride, rs);
}

// Column tiles per work-group: at most four, with the staged right-hand
// sides (rows padded to a multiple of 64) within 48 KB of LDS where possible.
static inline int plr_cap_nct(int n2)
{
    const int n2p = (n2 + 63) & ~63;
    for (int nct = 4; nct > 1; --nct)
        if ((size_t)n2p * (16 * nct + 1) * sizeof(double) <= 49152) return nct;
    return 1;
}

template <typename KT, typename VT, bool LP>
static int plr_cap_launch(const KT *Kinv, int d, int R, const VT *S, const VT *T, size_t tstride,
                          VT *W, size_t wstride, int trans, int ncols, float *wscale,
                          size_t sstride, hipStream_t st, size_t rs = 1)
{
    const int n2 = 2 * R, n2p = (n2 + 63) & ~63;
    if (ncols <= 0 || R <= 0) return 0;
    const int nct = plr_cap_nct(n2), cw = 16 * nct;
    const size_t lds = (size_t)n2p * (cw + 1) * sizeof(double);
    if (lds > 65536) {
        plr_set_error("rank too large for the capacitance kernel (LDS over 64 KB)");
        return -3;
    }
    dim3 grid(1u << d, (ncols + cw - 1) / cw);
#define PLR_CAP_CASE(n)                                                                            \
    case n:                                                                                        \
        hipLaunchKernelGGL((k_cap_mm<KT, VT, LP, n>), grid, dim3(256), lds, st, Kinv, R, S, T,     \
                           tstride, W, wstride, trans, ncols, wscale, sstride, rs);                \
        break;
    switch (nct) {
        PLR_CAP_CASE(1)
        PLR_CAP_CASE(2)
        PLR_CAP_CASE(3)
        PLR_CAP_CASE(4)
    }
#undef PLR_CAP_CASE
    PLR_HIP(hipGetLastError());
    return 0;
}
