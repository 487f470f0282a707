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
// column) also gets the power-of-two scale of its largest magnitude.
#pragma once

#include "plr_common.h"

typedef double plr_d4 __attribute__((ext_vector_type(4)));

template <typename KT, typename VT, bool LP>
__global__ __launch_bounds__(256) void k_cap_mm(const KT *__restrict__ Kinv, int R,
                                                const VT *__restrict__ S,
                                                const VT *__restrict__ T, size_t tstride,
                                                VT *__restrict__ W, size_t wstride, int trans,
                                                int ncols, int cw, float *__restrict__ wscale,
                                                size_t sstride)
{
    extern __shared__ double rhs[]; // [n2][cw]
    __shared__ unsigned smax[64][2];
    const int node = blockIdx.x, c0 = blockIdx.y * cw, n2 = 2 * R;
    const int nc = min(cw, ncols - c0);
    const int top = 2 * node, bot = top + 1;
    const KT *__restrict__ k = Kinv + (size_t)node * n2 * n2;
    for (int e = threadIdx.x; e < n2 * cw; e += blockDim.x) {
        const int j = e / cw, cl = e % cw;
        double v = 0.0;
        if (cl < nc) {
            const VT *t = T + (size_t)(c0 + cl) * tstride;
            if (!trans)
                v = (j < R) ? (double)t[(size_t)bot * R + j] : (double)t[(size_t)top * R + j - R];
            else
                v = (j < R) ? (double)S[(size_t)top * R + j] * (double)t[(size_t)top * R + j]
                            : (double)S[(size_t)bot * R + j - R] * (double)t[(size_t)bot * R + j - R];
        }
        rhs[e] = v;
    }
    if (LP)
        for (int e = threadIdx.x; e < 128; e += blockDim.x) smax[e >> 1][e & 1] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int ntile = (n2 + 15) / 16, nct = cw / 16;
    for (int it = wv; it < ntile; it += 4) {
        const int ia = min(it * 16 + m, n2 - 1);
        const bool arow = it * 16 + m < n2;
        plr_d4 acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = plr_d4{0.0, 0.0, 0.0, 0.0};
        for (int kb0 = 0; kb0 < n2; kb0 += 64) {
            double av[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int j = min(kb0 + 4 * u + g, n2 - 1);
                const KT kv = trans ? k[(size_t)ia * n2 + j] : k[(size_t)j * n2 + ia];
                av[u] = (arow && kb0 + 4 * u < n2) ? (double)kv : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                if (kb0 + 4 * u >= n2) break;
                const double *bq = rhs + (size_t)(kb0 + 4 * u + g) * cw + m;
#pragma unroll
                for (int ct = 0; ct < 4; ++ct)
                    if (ct < nct)
                        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bq[ct * 16], acc[ct], 0,
                                                                       0, 0);
            }
        }
        // lane holds c[it*16 + g + 4v][ct*16 + m]
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const int cl = ct * 16 + m;
            if (ct >= nct || cl >= nc) continue;
            VT *w = W + (size_t)(c0 + cl) * wstride;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = it * 16 + g + 4 * v;
                if (i >= n2) continue;
                const double c = acc[ct][v];
                if (!trans) {
                    const int sg = i < R ? top : bot, kk = i < R ? i : i - R;
                    const VT out = (VT)(-(double)S[(size_t)sg * R + kk] * c);
                    w[(size_t)sg * R + kk] = out;
                    if (LP) atomicMax(&smax[cl][i < R ? 0 : 1], __float_as_uint(fabsf((float)out)));
                } else {
                    const int sg = i < R ? bot : top, kk = i < R ? i : i - R;
                    w[(size_t)sg * R + kk] = (VT)(-c);
                }
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

// Columns per work-group: a multiple of 16, at most 64, with the staged
// right-hand sides within 48 KB of LDS (64 KB for the largest orders).
static inline int plr_cap_cw(int n2)
{
    int cw = 49152 / (n2 * (int)sizeof(double));
    cw = (cw / 16) * 16;
    return cw < 16 ? 16 : (cw > 64 ? 64 : cw);
}

template <typename KT, typename VT, bool LP>
static int plr_cap_launch(const KT *Kinv, int d, int R, const VT *S, const VT *T, size_t tstride,
                          VT *W, size_t wstride, int trans, int ncols, float *wscale,
                          size_t sstride, hipStream_t st)
{
    const int n2 = 2 * R;
    if (ncols <= 0 || R <= 0) return 0;
    const int cw = plr_cap_cw(n2);
    const size_t lds = (size_t)n2 * cw * sizeof(double);
    if (lds > 65536) {
        plr_set_error("rank too large for the capacitance kernel (2R x 16 doubles > 64 KB)");
        return -3;
    }
    dim3 grid(1u << d, (ncols + cw - 1) / cw);
    hipLaunchKernelGGL((k_cap_mm<KT, VT, LP>), grid, dim3(256), lds, st, Kinv, R, S, T, tstride, W,
                       wstride, trans, ncols, cw, wscale, sstride);
    PLR_HIP(hipGetLastError());
    return 0;
}
