// Single-precision building blocks of the multi-RHS products: storage FP32
// or BF16, arithmetic FP32 on v_mfma_f32_16x16x4_f32, right-hand sides and
// results FP32.  The task decomposition and the tables are those of the
// FP64 bodies (plr_mm.h: MmTab, mm_tab); a table built for these bodies
// carries operand pointers of the storage type in its P, Q slots and FP32
// sig in its S slots (sp_tab_ptr converts), which only these bodies read.
//
// MFMA layout of the 16x16x4 FP32 form, lane l (m = l % 16, g = l / 16):
//   A (16x4): A[m][g],  B (4x16): B[g][m],  D (16x16): D[4g + v][m], v = 0..3.
#pragma once

#include <cstdint>

#include "plr_mm.h"

typedef float sp_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ sp_f4 mfma32(float a, float b, sp_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// row of element v of a lane group's D fragment
__device__ __forceinline__ int sp_drow(int g, int v) { return 4 * g + v; }

// BF16 is stored as its 16 bits (uint16_t); a value widens to FP32 exactly.
__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float((unsigned)h << 16); }

template <typename T>
struct SpLd;
template <>
struct SpLd<float>
{
    static __device__ __forceinline__ float one(const float *p) { return *p; }
    // eight consecutive values (16-byte aligned)
    static __device__ __forceinline__ void eight(const float *p, float (&v)[8])
    {
        const float4 a = *(const float4 *)p, b = *(const float4 *)(p + 4);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
        v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    }
};
template <>
struct SpLd<uint16_t>
{
    static __device__ __forceinline__ float one(const uint16_t *p) { return bf16_f(*p); }
    static __device__ __forceinline__ void eight(const uint16_t *p, float (&v)[8])
    {
        const uint4 a = *(const uint4 *)p;
        const unsigned u[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[2 * j] = __uint_as_float(u[j] << 16);
            v[2 * j + 1] = __uint_as_float(u[j] & 0xffff0000u);
        }
    }
};

// T[c][s*R + k] (raw) or alpha * T to the partner segment, for one task of
// the table (one 64-lane wave: 32 ranks x 64 columns x up to T rows of one
// segment); long segments leave partial blocks [c_local][k_local] for
// sp_finish_body.  Rows are taken 16 per lane group (lane group g owns rows
// rb + 16g .. +15 of each 64-row block), eight loads in flight per operand.
template <typename T>
__device__ __forceinline__ void sp_reduce_body(const MmTab &a, int task, const float *__restrict__ X,
                                               int ldx, float alpha, int raw,
                                               float *__restrict__ part, float *__restrict__ w,
                                               size_t wstride)
{
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    int d = 0;
    while (task >= a.task_off[d + 1]) ++d;
    const int loc = task - a.task_off[d];
    const int nkg = a.nkg[d], ncg = a.ncg, tps = a.tps[d];
    const int kg = loc % nkg;
    const int cg = (loc / nkg) % ncg;
    const int st = loc / (nkg * ncg);
    const int s = st / tps, tile = st - s * tps;
    const int R = a.R[d], N = a.N, seg = a.seglen[d], Tt = a.T[d];
    const int k0 = kg * MM_KG, c0 = cg * MM_CG;
    const int rbeg = s * seg + tile * Tt;
    const int rend = min(rbeg + Tt, (s + 1) * seg);
    const int nct = min(4, (a.ncols - c0 + 15) >> 4); // live 16-column tiles (uniform)
    const T *P = (const T *)a.P[d];
    const T *pa[2];
    const float *pb[4];
    float va[2], vb[4]; // 1 for a real rank / column, 0 for padding
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
        const int k = k0 + kt * 16 + m;
        va[kt] = k < R ? 1.f : 0.f;
        pa[kt] = P + (size_t)(k < R ? k : 0) * N;
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int c = c0 + ct * 16 + m;
        vb[ct] = c < a.ncols ? 1.f : 0.f;
        pb[ct] = X + (size_t)(c < a.ncols ? c : 0) * ldx;
    }
    sp_f4 acc[2][4];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[kt][ct] = sp_f4{0.f, 0.f, 0.f, 0.f};
    for (int rb = rbeg; rb < rend; rb += 64) {
        const int r = rb + 16 * g;
#pragma unroll
        for (int t0 = 0; t0 < 16; t0 += 8) {
            float ua[2][8], ub[4][8];
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) SpLd<T>::eight(pa[kt] + r + t0, ua[kt]);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                if (ct >= nct) break;
                SpLd<float>::eight(pb[ct] + r + t0, ub[ct]);
            }
#pragma unroll
            for (int tt = 0; tt < 8; ++tt)
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    if (ct >= nct) break;
#pragma unroll
                    for (int kt = 0; kt < 2; ++kt)
                        acc[kt][ct] = mfma32(ua[kt][tt] * va[kt], ub[ct][tt] * vb[ct], acc[kt][ct]);
                }
        }
    }
    // lane holds T[k0 + 16kt + 4g + v][c0 + 16ct + m]
    if (tps == 1) {
        const int sp = raw ? s : (s ^ 1);
#pragma unroll
        for (int kt = 0; kt < 2; ++kt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int c = c0 + ct * 16 + m;
                if (c >= a.ncols) continue;
                float *wc = w + (size_t)c * wstride + a.w_off[d] + (size_t)sp * R;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = k0 + kt * 16 + sp_drow(g, v);
                    if (k < R) wc[k] = raw ? acc[kt][ct][v] : alpha * acc[kt][ct][v];
                }
            }
        return;
    }
    float *pp = part + ((size_t)a.part_off[d] + ((size_t)st * nkg + kg) * ncg + cg) * MM_BLK;
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                pp[(ct * 16 + m) * MM_KG + kt * 16 + sp_drow(g, v)] = acc[kt][ct][v];
}

// Sum of a segment's partial blocks (fixed order), one work-group of four
// waves per (level, segment, rank group, column group, 64-element chunk).
__device__ __forceinline__ void sp_finish_body(const MmTab &a, int bx, float alpha, int raw,
                                               const float *__restrict__ part,
                                               float *__restrict__ w, size_t wstride)
{
    __shared__ float red[4][64];
    const int task = bx / MM_FCH, chunk = bx % MM_FCH;
    int d = 0;
    while (task >= a.fin_off[d + 1]) ++d;
    const int loc = task - a.fin_off[d];
    const int nkg = a.nkg[d], ncg = a.ncg, tps = a.tps[d], R = a.R[d];
    const int kg = loc % nkg, cg = (loc / nkg) % ncg, s = loc / (nkg * ncg);
    const int el = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int e = chunk * 64 + el;
    const float *pp = part + ((size_t)a.part_off[d] + ((size_t)s * tps * nkg + kg) * ncg + cg) * MM_BLK + e;
    const size_t step = (size_t)nkg * ncg * MM_BLK;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j0 = ph; j0 < tps; j0 += 32) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = j0 + 4 * u;
            const float v = pp[(size_t)min(j, tps - 1) * step];
            acc[u] += j < tps ? v : 0.f;
        }
    }
    red[ph][el] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (ph != 0) return;
    const float t = (red[0][el] + red[1][el]) + (red[2][el] + red[3][el]);
    const int kl = e % MM_KG, cl = e / MM_KG;
    const int k = kg * MM_KG + kl, c = cg * MM_CG + cl;
    if (k >= R || c >= a.ncols) return;
    const int sp = raw ? s : (s ^ 1);
    w[(size_t)c * wstride + a.w_off[d] + (size_t)sp * R + k] = raw ? t : alpha * t;
}

// sh[c*65 + kk] = scale[kk] * src[c*cstride + kk] for c < nc, kk < kn, 0 for
// kk >= kn; rows of the 16-column tiles past nc are left as they are (they
// only feed output columns that are never stored)
__device__ __forceinline__ void sp_stage(float *sh, const float *src, size_t cstride, int nc,
                                         int kn, const float *scale = nullptr)
{
    const int ne = ((nc + 15) & ~15) * 64;
    for (int e0 = 0; e0 < ne; e0 += 8 * 256) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (e0 + u * 256 >= ne) break; // uniform
            const int e = e0 + u * 256 + (int)threadIdx.x, c = e >> 6, kk = e & 63;
            const float x = src[(size_t)min(c, nc - 1) * cstride + min(kk, kn - 1)];
            const float f = scale ? scale[min(kk, kn - 1)] : 1.f;
            v[u] = x * f * (float)(c < nc && kk < kn);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (e0 + u * 256 >= ne) break;
            const int e = e0 + u * 256 + (int)threadIdx.x;
            sh[(e >> 6) * 65 + (e & 63)] = v[u];
        }
    }
}

// Y^T (64 columns x 64 rows per work-group of four waves, 16 rows each) =
// beta Y^T + alpha X_leaf^T op(D)^T + sum over the table's levels of
// W_d^T (sig .* Q_d^T) (sig only for apply tables, a.sx).  X and W are
// staged in LDS per 64-wide chunk; D and Q come from HBM, 16 consecutive
// rows per lane group.
template <typename T, int NT = 4>
__device__ __forceinline__ void sp_expand_body(const MmTab &a, int bx, int by,
                                               const T *__restrict__ D,
                                               const float *__restrict__ X, int ldx, float alpha,
                                               float beta, int trans, const float *__restrict__ w,
                                               size_t wstride, float *__restrict__ Y, int ldy)
{
    __shared__ float sh[64 * 65];
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int wv = threadIdx.x >> 6;
    const int rb = bx * 64, r0 = rb + wv * 16;
    const int c0 = by * MM_CG;
    const int N = a.N, b = a.b, nc = min(MM_CG, a.ncols - c0);
    const int nct = min(NT, (nc + 15) >> 4); // live 16-column tiles (uniform)
    sp_f4 acc[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) acc[ct] = sp_f4{0.f, 0.f, 0.f, 0.f};
    // lane holds Y^T[c0 + 16ct + 4g + v][r0 + m]
    float yo[NT][4];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = min(c0 + ct * 16 + sp_drow(g, v), a.ncols - 1);
            yo[ct][v] = (ct < nct && beta != 0.f) ? Y[(size_t)c * ldy + r0 + m] : 0.f;
        }
    if (D) {
        const int leaf = rb / b, il = r0 - leaf * b;
        const size_t ds_i = trans ? b : 1, ds_j = trans ? 1 : b; // element (row, col) of op(D)
        const T *Dl = D + (size_t)leaf * b * b;
        const float *xl = X + (size_t)c0 * ldx + (size_t)leaf * b;
        for (int j0 = 0; j0 < b; j0 += 32) {
            const int jn = min(32, b - j0);
            float bq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int jj = j0 + min(4 * u, jn - 4) + g;
                bq[u] = SpLd<T>::one(Dl + (size_t)(il + m) * ds_i + (size_t)jj * ds_j);
            }
            __syncthreads();
            sp_stage(sh, xl + j0, ldx, nc, jn);
            __syncthreads();
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                if (ct >= nct) continue; // uniform
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc[ct] = mfma32(sh[(16 * ct + m) * 65 + 4 * u + g], bq[u], acc[ct]);
            }
        }
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) acc[ct] *= alpha;
    }
    for (int d = 0; d < a.L; ++d) {
        const int R = a.R[d];
        if (R == 0) continue;
        const int s = rb / a.seglen[d];
        const T *Qd = (const T *)a.Q[d];
        const float *Sd = (const float *)a.S[d];
        const float *wd = w + (size_t)c0 * wstride + a.w_off[d] + (size_t)s * R;
        for (int k0 = 0; k0 < R; k0 += 32) {
            const int kn = min(32, R - k0);
            float bq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int kk = k0 + min(4 * u, kn - 4) + g;
                bq[u] = SpLd<T>::one(Qd + (size_t)kk * N + r0 + m);
            }
            __syncthreads();
            sp_stage(sh, wd + k0, wstride, nc, kn,
                     a.sx ? Sd + (size_t)(trans ? (s ^ 1) : s) * R + k0 : nullptr);
            __syncthreads();
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                if (ct >= nct) continue; // uniform
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc[ct] = mfma32(sh[(16 * ct + m) * 65 + 4 * u + g], bq[u], acc[ct]);
            }
        }
    }
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = c0 + ct * 16 + sp_drow(g, v);
            if (c >= a.ncols) continue;
            Y[(size_t)c * ldy + r0 + m] = beta * yo[ct][v] + acc[ct][v];
        }
}

// Operand pointer of storage type T in a table slot typed for FP64.
template <typename T>
static inline const double *sp_tab_ptr(const T *p)
{
    return reinterpret_cast<const double *>(p);
}
