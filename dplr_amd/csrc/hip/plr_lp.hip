// Mixed-precision factored solve with FP64-residual iterative refinement.
//
// Storage: V (column-major, as the FP64 plan) and Ut (row-major per level,
// so that one MFMA operand is eight consecutive ranks of one row) in f16 or
// bf16; the leaves' inverses in the same 16-bit type, row-major per leaf;
// sig in FP32.  Ranks are padded to Rp = a multiple of 32 (the K step of
// v_mfma_f32_16x16x32_{f16,bf16}).  Capacitance inverses: FP64 on the top
// two levels, FP32 below.  Leaf LU and every capacitance LU run in FP64.
// Working vectors are FP32; the 16-bit MFMA operands are converted on load.
//
// MFMA layout (measured on gfx950) for the 16x16x32 forms, lane l:
//   A (16x32): A[l%16][8(l/16) + j],  B (32x16): B[8(l/16) + j][l%16],
//   D (16x16): D[4(l/16) + v][l%16], v = 0..3.
//
// The low-precision solve has the structure of the FP64 one (plr_solve.hip):
// z = Dinv r, then per level d = L-1 .. 0: t = V_d^T z per segment,
// c = inv(K) [t_bot; t_top], z -= Ut_d (sig .* c).  The expansion
// coefficients are scaled by a power of two per (segment, column) before the
// conversion to 16 bits, so tiny sig cannot underflow f16.  Refinement:
// x = 0, r = b; repeat { s = max|r| per column; x += s * solve_lp(r / s);
// r = b - A x in FP64 }.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "plr_cap.h"
#include "plr_common.h"

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

template <typename T>
struct Lp;
template <>
struct Lp<_Float16>
{
    typedef h8 v8;
    static __device__ __forceinline__ f4 mfma(h8 a, h8 b, f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Lp<__bf16>
{
    typedef b8 v8;
    static __device__ __forceinline__ f4 mfma(b8 a, b8 b, f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

template <typename T>
__device__ __forceinline__ typename Lp<T>::v8 cvt8(f8 x, float s)
{
    typename Lp<T>::v8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = (T)(x[j] * s);
    return r;
}

// Pair storage (type 3): a factor entry is hi + lo, two f16 values, of the
// entry times a power of two chosen per level so that the largest entry is
// about LP_PAIR_TOP (lo stays a normal f16 for all but the tiniest
// entries); activations are split the same way, times LP_PAIR_Z.  A
// product then takes three MFMAs (the lo x lo term, 2^-22 relative, is
// dropped) and carries about 22 bits: FP32-like solves.
static constexpr float LP_PAIR_TOP = 4096.f, LP_PAIR_Z = 256.f;

template <typename T>
__device__ __forceinline__ void split8(f8 x, float s, typename Lp<T>::v8 &hi, typename Lp<T>::v8 &lo)
{
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float v = x[j] * s;
        const T h = (T)v;
        hi[j] = h;
        lo[j] = (T)(v - (float)h);
    }
}

// acc + (ah + al)(bh + bl) without the al bl term
template <typename T>
__device__ __forceinline__ f4 mfma3(typename Lp<T>::v8 ah, typename Lp<T>::v8 al,
                                    typename Lp<T>::v8 bh, typename Lp<T>::v8 bl, f4 acc)
{
    acc = Lp<T>::mfma(al, bh, acc);
    acc = Lp<T>::mfma(ah, bl, acc);
    return Lp<T>::mfma(ah, bh, acc);
}

static constexpr int LP_KG = 32, LP_CG = 64, LP_TMAX = 512, LP_BLK = LP_KG * LP_CG;

struct LpTab
{
    int L, N, b, ncols, ncg;
    int R[PLR_MAXL], seglen[PLR_MAXL], T[PLR_MAXL], tps[PLR_MAXL], nkg[PLR_MAXL];
    int task_off[PLR_MAXL + 1], part_off[PLR_MAXL + 1], fin_off[PLR_MAXL + 1];
    const void *P[PLR_MAXL]; // 16-bit, column-major [Rp][N]
    const void *Q[PLR_MAXL]; // 16-bit, row-major [N][Rp]
};

// T[c][s*Rp + k] = sum over segment s of P[k][i] Z[c][i]   (raw, FP32)
template <typename T16>
__global__ __launch_bounds__(64) void k_lp_reduce(LpTab a, const float *__restrict__ Z, int ldz,
                                                  float *__restrict__ part, float *__restrict__ out,
                                                  size_t ostride)
{
    typedef typename Lp<T16>::v8 v8;
    const int lane = threadIdx.x, m = lane & 15, g = lane >> 4;
    const int task = blockIdx.x;
    int d = 0;
    while (task >= a.task_off[d + 1]) ++d;
    const int loc = task - a.task_off[d];
    const int nkg = a.nkg[d], ncg = a.ncg, tps = a.tps[d];
    const int kg = loc % nkg, cg = (loc / nkg) % ncg, st = loc / (nkg * ncg);
    const int s = st / tps, tile = st - s * tps;
    const int R = a.R[d], N = a.N, seg = a.seglen[d], Tt = a.T[d];
    const int k0 = kg * LP_KG, c0 = cg * LP_CG;
    const int rbeg = s * seg + tile * Tt, rend = min(rbeg + Tt, (s + 1) * seg);
    const T16 *P = (const T16 *)a.P[d];
    const T16 *pa[2];
    const float *pb[4];
    float vb[4]; // 1 for a real column, 0 for padding (multiplied, not selected)
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) pa[kt] = P + (size_t)(k0 + 16 * kt + m) * N;
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        const int c = c0 + 16 * ct + m;
        vb[ct] = c < a.ncols ? 1.f : 0.f;
        pb[ct] = Z + (size_t)(c < a.ncols ? c : 0) * ldz;
    }
    f4 acc[2][4];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[kt][ct] = f4{0.f, 0.f, 0.f, 0.f};
    // four 32-row steps per batch, every load of the batch issued first
    // (segments are whole multiples of 64 rows; a step past rend is masked)
    for (int rb = rbeg; rb < rend; rb += 128) {
        v8 av[4][2];
        f8 zv[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = min(rb + 32 * u, rend - 32) + 8 * g;
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) av[u][kt] = *(const v8 *)(pa[kt] + r);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) zv[u][ct] = *(const f8 *)(pb[ct] + r);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float on = rb + 32 * u < rend ? 1.f : 0.f;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const v8 bv = cvt8<T16>(zv[u][ct], vb[ct] * on);
#pragma unroll
                for (int kt = 0; kt < 2; ++kt) acc[kt][ct] = Lp<T16>::mfma(av[u][kt], bv, acc[kt][ct]);
            }
        }
    }
    // lane holds T[k0 + 16kt + 4g + v][c0 + 16ct + m]
    if (tps == 1) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const int c = c0 + 16 * ct + m;
            if (c >= a.ncols) continue;
            float *o = out + (size_t)c * ostride + (size_t)s * R;
#pragma unroll
            for (int kt = 0; kt < 2; ++kt)
#pragma unroll
                for (int v = 0; v < 4; ++v) o[k0 + 16 * kt + 4 * g + v] = acc[kt][ct][v];
        }
        return;
    }
    float *pp = part + ((size_t)a.part_off[d] + ((size_t)st * nkg + kg) * ncg + cg) * LP_BLK;
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) pp[(16 * ct + m) * LP_KG + 16 * kt + 4 * g + v] = acc[kt][ct][v];
}

// Sum of the tile partials (fixed order): one work-group per (level,
// segment, rank group, column group, 64-element chunk); four waves take every
// fourth tile, eight loads in flight, combined through LDS.
static constexpr int LP_FCH = LP_BLK / 64;

__global__ __launch_bounds__(256) void k_lp_finish(LpTab a, const float *__restrict__ part,
                                                   float *__restrict__ out, size_t ostride)
{
    __shared__ float red[4][64];
    const int task = blockIdx.x / LP_FCH, chunk = blockIdx.x % LP_FCH;
    int d = 0;
    while (task >= a.fin_off[d + 1]) ++d;
    const int loc = task - a.fin_off[d];
    const int nkg = a.nkg[d], ncg = a.ncg, tps = a.tps[d], R = a.R[d];
    const int kg = loc % nkg, cg = (loc / nkg) % ncg, s = loc / (nkg * ncg);
    const int el = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int e = chunk * 64 + el;
    const float *pp = part + ((size_t)a.part_off[d] + ((size_t)s * tps * nkg + kg) * ncg + cg) * LP_BLK + e;
    const size_t step = (size_t)nkg * ncg * LP_BLK;
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
    const int kl = e % LP_KG, cl = e / LP_KG;
    const int c = cg * LP_CG + cl;
    if (c >= a.ncols) return;
    out[(size_t)c * ostride + (size_t)s * R + kg * LP_KG + kl] = t;
}

// The same sums for a table of one level, the work-group shape following the
// level's tile count: P = tps/16 tile phases (a power of two, 1..8) times
// 256/P elements, sixteen loads in flight per thread, so every level is one
// round trip and a level of few tiles takes few work-groups.
__host__ __device__ inline int lpf_phases(int tps)
{
    int ph = 1;
    while (ph < 8 && 32 * ph <= tps) ph *= 2;
    return ph;
}

__global__ __launch_bounds__(256) void k_lp_finish1(LpTab a, const float *__restrict__ part,
                                                    float *__restrict__ out, size_t ostride)
{
    __shared__ float red[256];
    const int tps = a.tps[0], P = lpf_phases(tps), E = 256 / P, nb = LP_BLK / E;
    const int task = blockIdx.x / nb, blk = blockIdx.x % nb;
    const int nkg = a.nkg[0], ncg = a.ncg, R = a.R[0];
    const int kg = task % nkg, cg = (task / nkg) % ncg, s = task / (nkg * ncg);
    const int el = threadIdx.x % E, ph = threadIdx.x / E;
    const int e = blk * E + el;
    const float *pp = part + ((size_t)a.part_off[0] + ((size_t)s * tps * nkg + kg) * ncg + cg) * LP_BLK + e;
    const size_t step = (size_t)nkg * ncg * LP_BLK;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int j0 = ph; j0 < tps; j0 += 16 * P) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = pp[(size_t)min(j0 + P * u, tps - 1) * step];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u & 7] += j0 + P * u < tps ? v[u] : 0.f;
    }
    red[threadIdx.x] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (ph != 0) return;
    float t = red[el];
    for (int q = 1; q < P; ++q) t += red[q * E + el];
    const int kl = e % LP_KG, cl = e / LP_KG;
    const int c = cg * LP_CG + cl;
    if (c >= a.ncols) return;
    out[(size_t)c * ostride + (size_t)s * R + kg * LP_KG + kl] = t;
}

// Z^T (64 columns x 16 rows per wavefront): Z := beta*Z + Dinv X (leaf, when
// Dl != null) + sum over the table's levels of Q_d (wscale .* W_d).
template <typename T16>
__global__ __launch_bounds__(256) void k_lp_expand(LpTab a, const T16 *__restrict__ Dl,
                                                   const float *__restrict__ X, int ldx,
                                                   const float *__restrict__ W, size_t wstride,
                                                   const float *__restrict__ wscale, size_t sstride,
                                                   float beta, float *__restrict__ Z, int ldz)
{
    typedef typename Lp<T16>::v8 v8;
    const int lane = threadIdx.x & 63, m = lane & 15, g = lane >> 4;
    const int r0 = blockIdx.x * 64 + (threadIdx.x >> 6) * 16;
    const int c0 = blockIdx.y * LP_CG;
    const int N = a.N, b = a.b;
    if (r0 >= N) return;
    // Every load is unconditional (clamped index) and padding columns are
    // zeroed by a 0/1 factor, so the loads of one level go out together.
    int cc[4];
    float vc[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        cc[ct] = min(c0 + 16 * ct + m, a.ncols - 1);
        vc[ct] = c0 + 16 * ct + m < a.ncols ? 1.f : 0.f;
    }
    int cd[4][4]; // columns of the D rows this lane holds: c0 + 16ct + 4g + v
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) cd[ct][v] = min(c0 + 16 * ct + 4 * g + v, a.ncols - 1);
    f4 acc[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) acc[ct] = f4{0.f, 0.f, 0.f, 0.f};
    // the old Z values are loaded first (used only when beta != 0), so their
    // round trip overlaps the operand loads below
    float zo[4][4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) zo[ct][v] = Z[(size_t)cd[ct][v] * ldz + r0 + m];
    if (Dl) {
        const int leaf = r0 / b, il = r0 - leaf * b;
        const T16 *D = Dl + (size_t)leaf * b * b; // row-major
        for (int j0 = 0; j0 < b; j0 += 32) {
            const v8 bq = *(const v8 *)(D + (size_t)(il + m) * b + j0 + 8 * g);
            f8 x[4];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                x[ct] = *(const f8 *)(X + (size_t)cc[ct] * ldx + (size_t)leaf * b + j0 + 8 * g);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[ct] = Lp<T16>::mfma(cvt8<T16>(x[ct], vc[ct]), bq, acc[ct]);
        }
    }
    for (int d = 0; d < a.L; ++d) {
        const int R = a.R[d], s = r0 / a.seglen[d];
        const T16 *Q = (const T16 *)a.Q[d];
        const float *wd = W + (size_t)s * R;
        float isc[4], sc[4][4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) isc[ct] = wscale[(size_t)cc[ct] * sstride + s];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) sc[ct][v] = wscale[(size_t)cd[ct][v] * sstride + s];
        f4 al[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) al[ct] = f4{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < R; k0 += 32) {
            const v8 bq = *(const v8 *)(Q + (size_t)(r0 + m) * R + k0 + 8 * g);
            f8 w[4];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) w[ct] = *(const f8 *)(wd + (size_t)cc[ct] * wstride + k0 + 8 * g);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                al[ct] = Lp<T16>::mfma(cvt8<T16>(w[ct], vc[ct] / isc[ct]), bq, al[ct]);
        }
        // D row 4g+v is column c0 + 16ct + 4g + v: undo that column's scale
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[ct][v] += al[ct][v] * sc[ct][v];
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = c0 + 16 * ct + 4 * g + v;
            if (c < a.ncols)
                Z[(size_t)c * ldz + r0 + m] = (beta == 0.f ? 0.f : beta * zo[ct][v]) + acc[ct][v];
        }
}

// K = [I, M_bot S_bot; M_top S_top, I] in FP64 from the FP32 raw reduction.
__global__ void k_lp_capf_build(const float *T, size_t tstride, int off, int R, const float *S,
                                double *K)
{
    const int node = blockIdx.x, n2 = 2 * R;
    double *k = K + (size_t)node * n2 * n2;
    const int top = 2 * node, bot = top + 1;
    for (int e = threadIdx.x; e < n2 * n2; e += blockDim.x) {
        const int i = e % n2, j = e / n2;
        double v = (i == j) ? 1.0 : 0.0;
        if (i < R && j >= R)
            v += (double)T[(size_t)(off + j - R) * tstride + (size_t)bot * R + i] * S[(size_t)bot * R + j - R];
        else if (i >= R && j < R)
            v += (double)T[(size_t)(off + j) * tstride + (size_t)top * R + i - R] * S[(size_t)top * R + j];
        k[e] = v;
    }
}

// ---- conversions ------------------------------------------------------------

// dst[k][i] (column-major Rp x N) from the FP64 level array (R columns).
template <typename T16>
__global__ void k_lp_cvt_cols(const double *src, int R, int Rp, int N, T16 *dst)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)Rp * N) return;
    const int k = (int)(e / N);
    dst[e] = (T16)(k < R ? src[e] : 0.0);
}

// FP32 column-major copy with rank padding: dst[k][i] (Rp x N) from src (R x N).
__global__ void k_lp_f64_to_f32_cols(const double *src, int R, int Rp, int N, float *dst)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)Rp * N) return;
    const int k = (int)(e / N);
    dst[e] = k < R ? (float)src[e] : 0.f;
}

// Row-major 16-bit copy of an FP32 column-major level block (Rp x N) -> (N x Rp).
template <typename T16>
__global__ void k_lp_rowmajor(const float *src, int Rp, int N, T16 *dst)
{
    __shared__ float tile[32][33];
    const int i0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8)
        if (i0 + tx < N && k0 + j < Rp) tile[j][tx] = src[(size_t)(k0 + j) * N + i0 + tx];
    __syncthreads();
    for (int j = ty; j < 32; j += 8)
        if (i0 + j < N && k0 + tx < Rp) dst[(size_t)(i0 + j) * Rp + k0 + tx] = (T16)tile[tx][j];
}

// Row-major 16-bit leaves from column-major FP64 ones.
template <typename T16>
__global__ void k_lp_leaf_rowmajor(const double *src, int b, size_t nleaf, T16 *dst)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t bb = (size_t)b * b;
    if (e >= nleaf * bb) return;
    const size_t l = e / bb, r = e % bb;
    const int i = (int)(r / b), j = (int)(r % b); // dst (i, j) row-major
    dst[e] = (T16)src[l * bb + (size_t)j * b + i];
}

__global__ void k_lp_sig(const double *S, int R, int Rp, int nseg, float *Sf)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nseg * Rp) return;
    const int s = e / Rp, k = e % Rp;
    Sf[e] = k < R ? (float)S[(size_t)s * R + k] : 0.f;
}

// ---- pair storage: hi + lo f16 planes of s * x -----------------------------

__device__ __forceinline__ void pair_of(double v, _Float16 *hi, _Float16 *lo, size_t e)
{
    const _Float16 h = (_Float16)v;
    hi[e] = h;
    lo[e] = (_Float16)(v - (double)h);
}

// column-major Rp x N planes from an FP64 level array of R columns (ld N)
__global__ void k_lp_pair_cols(const double *src, int R, int Rp, int N, double s, _Float16 *hi,
                               _Float16 *lo)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)Rp * N) return;
    const int k = (int)(e / N);
    pair_of(k < R ? src[e] * s : 0.0, hi, lo, e);
}

// row-major N x Rp planes from an FP64 level array of R columns (ld N)
__global__ void k_lp_pair_rows(const double *src, int R, int Rp, int N, double s, _Float16 *hi,
                               _Float16 *lo)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)Rp * N) return;
    const size_t i = e / Rp;
    const int k = (int)(e % Rp);
    pair_of(k < R ? src[(size_t)k * N + i] * s : 0.0, hi, lo, e);
}

// row-major leaf planes from column-major FP64 leaves
__global__ void k_lp_pair_leaf(const double *src, int b, size_t nleaf, double s, _Float16 *hi,
                               _Float16 *lo)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t bb = (size_t)b * b;
    if (e >= nleaf * bb) return;
    const size_t l = e / bb, r = e % bb;
    const int i = (int)(r / b), j = (int)(r % b);
    pair_of(src[l * bb + (size_t)j * b + i] * s, hi, lo, e);
}

// largest |x| of n entries, as float bits (unsigned order = numeric
// order), folded into *out: a wave maximum, the block's four through LDS,
// one atomic per block
__global__ __launch_bounds__(256) void k_lp_maxabs(const double *x, size_t n, unsigned *out)
{
    __shared__ unsigned red[4];
    unsigned mv = 0u;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x)
        mv = max(mv, __float_as_uint(fabsf((float)x[e])));
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) mv = max(mv, (unsigned)__shfl_xor((int)mv, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mv;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, max(max(red[0], red[1]), max(red[2], red[3])));
}

// inv(K) of every node with its ranks padded from R to Rp: K's index i < R
// (bottom segment) stays, i >= R (top) moves to i - R + Rp; padded rows and
// columns are those of the identity (their factor columns are zero).
template <typename TK>
__global__ void k_lp_kpad(const double *src, int R, int Rp, TK *dst)
{
    const int node = blockIdx.x, n2 = 2 * R, n2p = 2 * Rp;
    const double *s = src + (size_t)node * n2 * n2;
    TK *d = dst + (size_t)node * n2p * n2p;
    for (int e = threadIdx.x; e < n2p * n2p; e += blockDim.x) {
        const int ip = e % n2p, jp = e / n2p;
        const int i = ip < Rp ? (ip < R ? ip : -1) : (ip - Rp < R ? ip - Rp + R : -1);
        const int j = jp < Rp ? (jp < R ? jp : -1) : (jp - Rp < R ? jp - Rp + R : -1);
        d[e] = (TK)((i < 0 || j < 0) ? (ip == jp ? 1.0 : 0.0) : s[(size_t)j * n2 + i]);
    }
}

__global__ void k_f64_to_f32(const double *src, size_t n, float *dst)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = (float)src[e];
}

// Per column: scale[c] = max |r[c]|, r32 = r / scale (1 if zero).  Two
// kernels: partial maxima of 64 row chunks per column, then the scaling
// (each block folds the 64 partials itself).
static constexpr int LP_SCALE_CHUNKS = 64;

__global__ __launch_bounds__(256) void k_lp_colmax(const double *r, int ldr, int N, double *pmax)
{
    __shared__ double red[4];
    const int c = blockIdx.y, ch = blockIdx.x;
    const int len = (N + LP_SCALE_CHUNKS - 1) / LP_SCALE_CHUNKS;
    const int beg = ch * len, end = min(N, beg + len);
    const double *rc = r + (size_t)c * ldr;
    double mv = 0.0;
    if ((len & 3) == 0 && (N & 3) == 0 && ((uintptr_t)rc & 15) == 0) {
        // four rows per thread as two 16-byte loads, both in flight at once
        for (int i = beg + 4 * (int)threadIdx.x; i < end; i += 1024) {
            const double2 a0 = *(const double2 *)(rc + i), a1 = *(const double2 *)(rc + i + 2);
            mv = fmax(mv, fmax(fmax(fabs(a0.x), fabs(a0.y)), fmax(fabs(a1.x), fabs(a1.y))));
        }
    } else {
        for (int i = beg + threadIdx.x; i < end; i += 256) mv = fmax(mv, fabs(rc[i]));
    }
    // the wave's maximum, then the four waves' through LDS
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) mv = fmax(mv, __shfl_xor(mv, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mv;
    __syncthreads();
    if (threadIdx.x == 0)
        pmax[(size_t)c * LP_SCALE_CHUNKS + ch] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Copies of the correction-norm sums (see k_lp_update), and the most
// refinement steps whose sums are kept at once (one slot per step).
static constexpr int LP_NRM_COPIES = 16, LP_NRM_SLOTS = 16;

// pm: pm_n partial maxima of |r| per column (column stride pm_n).
__global__ __launch_bounds__(256) void k_lp_scale_rhs(const double *r, int ldr, int N,
                                                      const double *pm, int pm_n, double *scale,
                                                      float *r32, double *nrm)
{
    const int c = blockIdx.y;
    if (nrm && blockIdx.x == 0 && threadIdx.x < 2 * LP_NRM_COPIES) // norms of this step: zero
        nrm[((size_t)(threadIdx.x >> 1) * gridDim.y + c) * 2 + (threadIdx.x & 1)] = 0.0;
    double mv = 0.0;
    for (int j = 0; j < pm_n; ++j) mv = fmax(mv, pm[(size_t)c * pm_n + j]);
    if (blockIdx.x == 0 && threadIdx.x == 0) scale[c] = mv;
    const double inv = 1.0 / (mv > 0.0 ? mv : 1.0);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N) r32[(size_t)c * N + i] = (float)(r[(size_t)c * ldr + i] * inv);
}

// out[c] = max_j pm[c * pm_n + j]
__global__ void k_lp_pmfold(const double *pm, int pm_n, int n, double *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double mv = 0.0;
    for (int j = 0; j < pm_n; ++j) mv = fmax(mv, pm[(size_t)c * pm_n + j]);
    out[c] = mv;
}

// Norms of a correction step, for the refinement's error estimate: per
// column the sums of dx^2 and of x^2 after the update (FP64).  Kernels add
// their work-group's sums into copy (blockIdx.x % LP_NRM_COPIES) of
// nrm[copy][column][2], spreading the atomics over several cache lines; the
// host adds the copies.  NaN and infinity propagate through the sums.


// x[c] (+)= scale[c] * z[c]  (FP64 accumulation of the correction; the
// first correction sets x); grid (row blocks of 256, columns).
__global__ __launch_bounds__(256) void k_lp_update(const float *z, int N, const double *scale,
                                                   double *x, int ldx, int n, int first,
                                                   double *nrm)
{
    const int c = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
    double dd = 0.0, xx = 0.0;
    if (i < N) {
        const double sc = scale[c] > 0.0 ? scale[c] : 1.0;
        double *xe = x + (size_t)c * ldx + i;
        const double d = sc * (double)z[(size_t)c * N + i];
        const double xn = (first ? 0.0 : *xe) + d;
        *xe = xn;
        dd = d * d;
        xx = xn * xn;
    }
    if (nrm) {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            dd += __shfl_xor(dd, o, 64);
            xx += __shfl_xor(xx, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            double *q = nrm + ((size_t)(blockIdx.x % LP_NRM_COPIES) * n + c) * 2;
            atomicAdd(q, dd);
            atomicAdd(q + 1, xx);
        }
    }
}

// ---------------------------------------------------------------------------
// host side

struct plr_lp
{
    int type; // 1 = f16, 2 = bf16, 3 = f16 pairs (hi + lo planes)
    float isV[PLR_MAXL], isU[PLR_MAXL], isD; // pairs: inverse factor scales (1 otherwise)
    int Rp[PLR_MAXL];
    size_t off[PLR_MAXL + 1];  // element offset of level d in Vh / Uh / Ut32 (Rp columns of N)
    size_t soff[PLR_MAXL + 1]; // element offset of level d in Sf
    size_t koff[PLR_MAXL + 1]; // element offset of level d in the K inverses
    void *Vh, *Uh, *Dh;        // 16-bit factors
    float *Sf;
    double *K64;               // FP64 capacitance inverses (top two levels)
    float *K32;                // FP32 capacitance inverses (other levels)
    // solve workspace
    int cols;
    size_t t_per_col, s_per_col;
    float *T, *W, *wscale, *part, *r32, *z32;
    size_t part_cap;
    double *r, *scale, *pmax;
    double *bmax; // maxima of |r| per (column, 128-row block) from the residual product
    double *nrm;      // norms of the correction steps (lp_correct): per step a slot of
                      // LP_NRM_COPIES x cols x 2, LP_NRM_SLOTS slots
    double *nrm_host; // pinned host copy of nrm
    int pred;         // solves the last refined solve to a tolerance needed,
    double pred_tol;  // and that tolerance
};

static void lp_free(plr_lp *q)
{
    if (!q) return;
    void *b[] = {q->Vh, q->Uh, q->Dh, q->Sf, q->K64, q->K32, q->T, q->W, q->wscale, q->part,
                 q->r32, q->z32, q->r, q->scale, q->pmax, q->bmax, q->nrm};
    for (void *x : b)
        if (x) (void)hipFree(x);
    if (q->nrm_host) (void)hipHostFree(q->nrm_host);
    delete q;
}

struct lp_entry
{
    const dplrh_plan *p;
    plr_lp *q;
};
static std::vector<lp_entry> g_lp;

static plr_lp *lp_of(const dplrh_plan *p)
{
    for (auto &e : g_lp)
        if (e.p == p) return e.q;
    return nullptr;
}

void plr_lp_release(const dplrh_plan *p)
{
    for (size_t i = 0; i < g_lp.size(); ++i)
        if (g_lp[i].p == p) {
            lp_free(g_lp[i].q);
            g_lp.erase(g_lp.begin() + i);
            return;
        }
}

static void lp_tab(const dplrh_plan *p, const plr_lp *q, int d0, int d1, int ncols, LpTab *t,
                   int tmax = LP_TMAX)
{
    memset(t, 0, sizeof(*t));
    t->L = d1 - d0;
    t->N = p->N;
    t->b = p->b;
    t->ncols = ncols;
    t->ncg = (ncols + LP_CG - 1) / LP_CG;
    const size_t es = 2; // bytes per 16-bit element
    int task = 0, part = 0, fin = 0;
    for (int i = 0; i < t->L; ++i) {
        const int d = d0 + i, seg = p->N >> (d + 1), nseg = 2 << d, R = q->Rp[d];
        const int T = seg < tmax ? seg : tmax, tps = seg / T, nkg = R / LP_KG;
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
        t->P[i] = (const char *)q->Vh + q->off[d] * es;
        t->Q[i] = (const char *)q->Uh + q->off[d] * es;
    }
    t->task_off[t->L] = task;
    t->part_off[t->L] = part;
    t->fin_off[t->L] = fin;
}

static int lp_part(const dplrh_plan *p, plr_lp *q, size_t blocks)
{
    const size_t need = blocks * LP_BLK;
    if (need <= q->part_cap) return 0;
    ++const_cast<dplrh_plan *>(p)->ws_gen; // buffers move: cached graphs are stale
    if (q->part) (void)hipFree(q->part);
    q->part = nullptr;
    q->part_cap = 0;
    PLR_HIP(hipMalloc(&q->part, need * sizeof(float)));
    q->part_cap = need;
    return 0;
}

static int lp_ws(const dplrh_plan *p, plr_lp *q, int cols)
{
    if (cols <= q->cols) return 0;
    ++const_cast<dplrh_plan *>(p)->ws_gen;
    void *old[] = {q->T, q->W, q->wscale, q->r32, q->z32, q->r, q->scale, q->pmax, q->nrm};
    for (void *x : old)
        if (x) (void)hipFree(x);
    if (q->nrm_host) (void)hipHostFree(q->nrm_host);
    q->T = q->W = q->wscale = q->r32 = q->z32 = nullptr;
    q->r = q->scale = q->pmax = nullptr;
    q->nrm = q->nrm_host = nullptr;
    q->cols = 0;
    size_t t = 32, s = 2;
    for (int d = 0; d < p->L; ++d) {
        const size_t nseg = (size_t)2 << d;
        if (nseg * q->Rp[d] > t) t = nseg * q->Rp[d];
        if (nseg > s) s = nseg;
    }
    q->t_per_col = (t + 7) & ~(size_t)7;
    q->s_per_col = s;
    PLR_HIP(hipMalloc(&q->T, q->t_per_col * cols * sizeof(float)));
    PLR_HIP(hipMalloc(&q->W, q->t_per_col * cols * sizeof(float)));
    PLR_HIP(hipMalloc(&q->wscale, q->s_per_col * cols * sizeof(float)));
    PLR_HIP(hipMalloc(&q->r32, (size_t)p->N * cols * sizeof(float)));
    PLR_HIP(hipMalloc(&q->z32, (size_t)p->N * cols * sizeof(float)));
    PLR_HIP(hipMalloc(&q->r, (size_t)p->N * cols * sizeof(double)));
    PLR_HIP(hipMalloc(&q->scale, (size_t)cols * sizeof(double)));
    PLR_HIP(hipMalloc(&q->pmax, (size_t)cols * LP_SCALE_CHUNKS * sizeof(double)));
    const size_t nn = (size_t)LP_NRM_SLOTS * LP_NRM_COPIES * 2 * cols;
    PLR_HIP(hipMalloc(&q->nrm, nn * sizeof(double)));
    PLR_HIP(hipHostMalloc((void **)&q->nrm_host, nn * sizeof(double), hipHostMallocDefault));
    q->cols = cols;
    return 0;
}

template <typename T16>
static int lp_reduce(const dplrh_plan *p, plr_lp *q, int d, const float *Z, int ldz, int ncols,
                     float *out, size_t ostride, hipStream_t st)
{
    LpTab t;
    lp_tab(p, q, d, d + 1, ncols, &t);
    int rc = lp_part(p, q, (size_t)t.part_off[1]);
    if (rc) return rc;
    if (t.task_off[1] > 0) {
        hipLaunchKernelGGL(k_lp_reduce<T16>, dim3(t.task_off[1]), dim3(64), 0, st, t, Z, ldz,
                           q->part, out, ostride);
        PLR_HIP(hipGetLastError());
    }
    if (t.fin_off[1] > 0) {
        hipLaunchKernelGGL(k_lp_finish, dim3(t.fin_off[1] * LP_FCH), dim3(256), 0, st, t, q->part, out,
                           ostride);
        PLR_HIP(hipGetLastError());
    }
    return 0;
}

template <typename T16>
static int lp_expand(const dplrh_plan *p, plr_lp *q, int d, const T16 *Dl, const float *X,
                     int ldx, const float *W, size_t wstride, const float *wscale, float beta,
                     float *Z, int ldz, int ncols, hipStream_t st)
{
    LpTab t;
    if (d >= 0)
        lp_tab(p, q, d, d + 1, ncols, &t);
    else
        lp_tab(p, q, 0, 0, ncols, &t);
    dim3 grid((p->N + 63) / 64, t.ncg);
    hipLaunchKernelGGL(k_lp_expand<T16>, grid, dim3(256), 0, st, t, Dl, X, ldx, W, wstride,
                       wscale, q->s_per_col, beta, Z, ldz);
    PLR_HIP(hipGetLastError());
    return 0;
}

template <typename T16>
static int lp_cap(const dplrh_plan *p, plr_lp *q, int d, const float *T, size_t tstride,
                  float *W, int ncols, hipStream_t st)
{
    (void)p;
    const int Rp = q->Rp[d];
    const float *Sf = q->Sf + q->soff[d];
    if (d < 2)
        return plr_cap_launch<double, float, true>(q->K64 + q->koff[d], d, Rp, Sf, T, tstride, W,
                                                   tstride, 0, ncols, q->wscale, q->s_per_col, st);
    return plr_cap_launch<float, float, true>(q->K32 + q->koff[d], d, Rp, Sf, T, tstride, W, tstride,
                                              0, ncols, q->wscale, q->s_per_col, st);
}


template <typename T16>
static int lp_factor_t(dplrh_plan *p, plr_lp *q, int *info, hipStream_t st)
{
    const int N = p->N, L = p->L, b = p->b, nleaf = N / b;
    // 1. leaves: FP64 LU + inverse, stored 16-bit row-major
    double *dinv = nullptr;
    PLR_HIP(hipMalloc(&dinv, (size_t)N * b * sizeof(double)));
    PLR_HIP(hipMemcpyAsync(dinv, p->D, (size_t)N * b * sizeof(double), hipMemcpyDeviceToDevice, st));
    double *dinv2 = nullptr;
    PLR_HIP(hipMalloc(&dinv2, (size_t)N * b * sizeof(double)));
    int rc = dplrh_lp_lu_inverse(dinv, b, nleaf, dinv2, info, st);
    if (rc || *info) {
        (void)hipFree(dinv);
        (void)hipFree(dinv2);
        return rc;
    }
    {
        const size_t n = (size_t)N * b;
        hipLaunchKernelGGL(k_lp_leaf_rowmajor<T16>, dim3((n + 255) / 256), dim3(256), 0, st, dinv2,
                           b, (size_t)nleaf, (T16 *)q->Dh);
        PLR_HIP(hipGetLastError());
    }
    // 2. V and sig in low precision; U in FP32 with padded ranks
    const size_t ctot = q->off[L] / (N ? N : 1);
    float *uw = nullptr, *ut = nullptr;
    PLR_HIP(hipMalloc(&uw, (q->off[L] ? q->off[L] : 1) * sizeof(float)));
    PLR_HIP(hipMalloc(&ut, (q->off[L] ? q->off[L] : 1) * sizeof(float)));
    for (int d = 0; d < L; ++d) {
        const size_t n = (size_t)q->Rp[d] * N;
        hipLaunchKernelGGL(k_lp_cvt_cols<T16>, dim3((n + 255) / 256), dim3(256), 0, st,
                           p->V + p->uoff[d], p->R[d], q->Rp[d], N, (T16 *)q->Vh + q->off[d]);
        hipLaunchKernelGGL(k_lp_f64_to_f32_cols, dim3((n + 255) / 256), dim3(256), 0, st,
                           p->U + p->uoff[d], p->R[d], q->Rp[d], N, uw + q->off[d]);
        const int ns = (2 << d) * q->Rp[d];
        hipLaunchKernelGGL(k_lp_sig, dim3((ns + 255) / 256), dim3(256), 0, st, p->S + p->soff[d],
                           p->R[d], q->Rp[d], 2 << d, q->Sf + q->soff[d]);
        PLR_HIP(hipGetLastError());
    }
    // 3. Ut = Dinv U
    if (ctot > 0) {
        rc = lp_expand<T16>(p, q, -1, (const T16 *)q->Dh, uw, N, nullptr, 0, nullptr, 0.f, ut, N,
                            (int)ctot, st);
        if (rc) return rc;
    }
    // 4. levels bottom-up
    size_t kmax = 1;
    for (int d = 0; d < L; ++d) {
        const size_t n2 = 2 * (size_t)q->Rp[d];
        if (((size_t)1 << d) * n2 * n2 > kmax) kmax = ((size_t)1 << d) * n2 * n2;
    }
    double *kbuf = nullptr, *kinv = nullptr;
    PLR_HIP(hipMalloc(&kbuf, kmax * sizeof(double)));
    PLR_HIP(hipMalloc(&kinv, kmax * sizeof(double)));
    rc = lp_ws(p, q, (int)(ctot ? ctot : 1));
    for (int d = L - 1; d >= 0 && !rc; --d) {
        const int Rp = q->Rp[d], n2 = 2 * Rp, nodes = 1 << d;
        const int off = (int)(q->off[d] / N);
        const size_t n = (size_t)Rp * N;
        dim3 g2((N + 31) / 32, (Rp + 31) / 32);
        hipLaunchKernelGGL(k_lp_rowmajor<T16>, g2, dim3(256), 0, st, ut + q->off[d], Rp, N,
                           (T16 *)q->Uh + q->off[d]);
        PLR_HIP(hipGetLastError());
        (void)n;
        if ((rc = lp_reduce<T16>(p, q, d, ut, N, off + Rp, q->T, q->t_per_col, st))) break;
        hipLaunchKernelGGL(k_lp_capf_build, dim3(nodes), dim3(256), 0, st, q->T, q->t_per_col, off,
                           Rp, q->Sf + q->soff[d], kbuf);
        PLR_HIP(hipGetLastError());
        int kinfo = 0;
        if ((rc = dplrh_lp_lu_inverse(kbuf, n2, nodes, kinv, &kinfo, st))) break;
        if (kinfo && !*info) *info = kinfo;
        const size_t nk = (size_t)nodes * n2 * n2;
        if (d < 2) {
            PLR_HIP(hipMemcpyAsync(q->K64 + q->koff[d], kinv, nk * sizeof(double),
                                   hipMemcpyDeviceToDevice, st));
        } else {
            hipLaunchKernelGGL(k_f64_to_f32, dim3((nk + 255) / 256), dim3(256), 0, st, kinv, nk,
                               q->K32 + q->koff[d]);
            PLR_HIP(hipGetLastError());
        }
        if (off > 0) {
            if ((rc = lp_cap<T16>(p, q, d, q->T, q->t_per_col, q->W, off, st))) break;
            if ((rc = lp_expand<T16>(p, q, d, nullptr, nullptr, 0, q->W, q->t_per_col, q->wscale,
                                     1.f, ut, N, off, st)))
                break;
        }
    }
    PLR_HIP(hipStreamSynchronize(st));
    (void)hipFree(dinv);
    (void)hipFree(dinv2);
    (void)hipFree(uw);
    (void)hipFree(ut);
    (void)hipFree(kbuf);
    (void)hipFree(kinv);
    return rc;
}

static bool lp_v2_ok(const dplrh_plan *p, const plr_lp *q);

// Pair storage (type 3) from the FP64 factorization, which runs first when
// the plan has none: V, Ut and the leaf inverses as hi + lo f16 planes of
// their FP64 values times a power of two per level (the largest entry at
// about LP_PAIR_TOP), sig in FP32, the capacitance inverses padded to Rp
// (FP64 on the top two levels, FP32 below).  The lo planes sit right after
// the hi planes (offset off[L] for V and Ut, N * b for the leaves).
static int lp_factor_pair(dplrh_plan *p, plr_lp *q, int *info, hipStream_t st)
{
    if (!p->f.done) {
        const int rc = dplrh_factor(p, info, st);
        if (rc || *info) return rc;
    }
    const int N = p->N, L = p->L, b = p->b;
    // largest |x| of V_d, Ut_d (every level) and the leaf inverses, read
    // back at once; slot 2d: V_d, 2d + 1: Ut_d, 2L: leaves
    const int nslot = 2 * L + 1;
    unsigned *dmax = nullptr;
    PLR_HIP(hipMalloc(&dmax, nslot * sizeof(unsigned)));
    PLR_HIP(hipMemsetAsync(dmax, 0, nslot * sizeof(unsigned), st));
    auto max_of = [&](const double *x, size_t n, int slot) -> int {
        if (n == 0) return 0;
        const size_t blocks = (n + 255) / 256 < 256 ? (n + 255) / 256 : 256;
        hipLaunchKernelGGL(k_lp_maxabs, dim3((unsigned)blocks), dim3(256), 0, st, x, n, dmax + slot);
        PLR_HIP(hipGetLastError());
        return 0;
    };
    int rc = 0;
    for (int d = 0; d < L && !rc; ++d) {
        const size_t n = (size_t)N * p->R[d];
        if (!(rc = max_of(p->V + p->uoff[d], n, 2 * d))) rc = max_of(p->f.Ut + p->uoff[d], n, 2 * d + 1);
    }
    if (!rc) rc = max_of(p->f.Dinv, (size_t)N * b, 2 * L);
    std::vector<unsigned> hmax(nslot, 0u);
    if (!rc) {
        PLR_HIP(hipMemcpyAsync(hmax.data(), dmax, nslot * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        PLR_HIP(hipStreamSynchronize(st));
    }
    // the power of two s with s * max|x| in (LP_PAIR_TOP / 2, LP_PAIR_TOP]
    auto scale_of = [&](int slot, double *s) -> int {
        *s = 1.0;
        float mx;
        memcpy(&mx, &hmax[slot], sizeof(mx));
        if (mx > 0.f && mx < INFINITY) {
            int e;
            (void)frexp((double)LP_PAIR_TOP / (double)mx, &e);
            *s = ldexp(1.0, e - 1);
        }
        return 0;
    };
    _Float16 *Vhi = (_Float16 *)q->Vh, *Vlo = Vhi + q->off[L];
    _Float16 *Uhi = (_Float16 *)q->Uh, *Ulo = Uhi + q->off[L];
    _Float16 *Dhi = (_Float16 *)q->Dh, *Dlo = Dhi + (size_t)N * b;
    for (int d = 0; d < L && !rc; ++d) {
        const int R = p->R[d], Rp = q->Rp[d];
        const size_t n = (size_t)N * R, np = (size_t)N * Rp;
        const unsigned gp = (unsigned)((np + 255) / 256);
        double sV = 1.0, sU = 1.0;
        (void)n;
        scale_of(2 * d, &sV);
        scale_of(2 * d + 1, &sU);
        if (np > 0) {
            hipLaunchKernelGGL(k_lp_pair_cols, dim3(gp), dim3(256), 0, st, p->V + p->uoff[d], R, Rp,
                               N, sV, Vhi + q->off[d], Vlo + q->off[d]);
            hipLaunchKernelGGL(k_lp_pair_rows, dim3(gp), dim3(256), 0, st, p->f.Ut + p->uoff[d], R,
                               Rp, N, sU, Uhi + q->off[d], Ulo + q->off[d]);
        }
        const int ns = (2 << d) * Rp;
        if (ns > 0)
            hipLaunchKernelGGL(k_lp_sig, dim3((ns + 255) / 256), dim3(256), 0, st, p->S + p->soff[d],
                               R, Rp, 2 << d, q->Sf + q->soff[d]);
        if (Rp > 0) {
            if (d < 2)
                hipLaunchKernelGGL(k_lp_kpad<double>, dim3(1 << d), dim3(256), 0, st,
                                   p->f.Kinv + p->f.koff[d], R, Rp, q->K64 + q->koff[d]);
            else
                hipLaunchKernelGGL(k_lp_kpad<float>, dim3(1 << d), dim3(256), 0, st,
                                   p->f.Kinv + p->f.koff[d], R, Rp, q->K32 + q->koff[d]);
        }
        if (hipGetLastError() != hipSuccess) {
            plr_set_error("pair factors: conversion launch failed");
            rc = -10;
        }
        q->isV[d] = (float)(1.0 / sV);
        q->isU[d] = (float)(1.0 / sU);
    }
    if (!rc) {
        double sD = 1.0;
        const size_t nd = (size_t)N * b;
        rc = scale_of(2 * L, &sD);
        if (!rc) {
            hipLaunchKernelGGL(k_lp_pair_leaf, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st,
                               p->f.Dinv, b, (size_t)(N / b), sD, Dhi, Dlo);
            q->isD = (float)(1.0 / sD);
            if (hipGetLastError() != hipSuccess) {
                plr_set_error("pair factors: leaf conversion launch failed");
                rc = -10;
            }
        }
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = -10;
    (void)hipFree(dmax);
    return rc;
}

// type: 1 = f16, 2 = bf16, 3 = f16 pairs.  Requires plr_mm_ok(p).
extern "C" int dplrh_lp_factor(dplrh_plan *p, int type, int *info, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    *info = 0;
    if (!plr_mm_ok(p)) {
        plr_set_error("mixed precision needs the leaf order to be a multiple of 64");
        return -5;
    }
    if (type != 1 && type != 2 && type != 3) {
        plr_set_error("unknown low-precision type %d", type);
        return -5;
    }
    plr_lp_release(p);
    ++p->ws_gen;
    plr_lp *q = new plr_lp;
    memset(q, 0, sizeof(*q));
    q->type = type;
    for (int d = 0; d < PLR_MAXL; ++d) q->isV[d] = q->isU[d] = 1.f;
    q->isD = 1.f;
    const size_t planes = type == 3 ? 2 : 1;
    size_t off = 0, so = 0, k64 = 0, k32 = 0;
    for (int d = 0; d < p->L; ++d) {
        q->Rp[d] = (p->R[d] + 31) & ~31;
        q->off[d] = off;
        off += (size_t)q->Rp[d] * p->N;
        q->soff[d] = so;
        so += ((size_t)2 << d) * q->Rp[d];
        const size_t nk = ((size_t)1 << d) * 4 * q->Rp[d] * q->Rp[d];
        q->koff[d] = d < 2 ? k64 : k32;
        if (d < 2)
            k64 += nk;
        else
            k32 += nk;
    }
    q->off[p->L] = off;
    q->soff[p->L] = so;
    int rc = 0;
    auto alloc = [&](void **ptr, size_t bytes) -> int {
        PLR_HIP(hipMalloc(ptr, bytes ? bytes : 8));
        return 0;
    };
    if ((rc = alloc(&q->Vh, off * 2 * planes)) || (rc = alloc(&q->Uh, off * 2 * planes)) ||
        (rc = alloc(&q->Dh, (size_t)p->N * p->b * 2 * planes)) || (rc = alloc((void **)&q->Sf, so * 4)) ||
        (rc = alloc((void **)&q->K64, k64 * 8)) || (rc = alloc((void **)&q->K32, k32 * 4))) {
        lp_free(q);
        return rc;
    }
    g_lp.push_back({p, q});
    rc = type == 1   ? lp_factor_t<_Float16>(p, q, info, st)
         : type == 2 ? lp_factor_t<__bf16>(p, q, info, st)
                     : lp_factor_pair(p, q, info, st);
    if (!rc && !*info && type == 3 && !lp_v2_ok(p, q)) {
        plr_set_error("f16 pair storage needs leaves of 64..512 rows, ranks <= 64 and a shape the "
                      "span kernels cover");
        rc = -5;
    }
    if (rc || *info) plr_lp_release(p);
    return rc;
}

template <typename T16>
static int lp_solve_t(const dplrh_plan *p, plr_lp *q, int n, hipStream_t st)
{
    const int N = p->N, L = p->L;
    int rc = lp_expand<T16>(p, q, -1, (const T16 *)q->Dh, q->r32, N, nullptr, 0, nullptr, 0.f,
                            q->z32, N, n, st);
    for (int d = L - 1; d >= 0 && !rc; --d) {
        if ((rc = lp_reduce<T16>(p, q, d, q->z32, N, n, q->T, q->t_per_col, st))) break;
        if ((rc = lp_cap<T16>(p, q, d, q->T, q->t_per_col, q->W, n, st))) break;
        rc = lp_expand<T16>(p, q, d, nullptr, nullptr, 0, q->W, q->t_per_col, q->wscale, 1.f,
                            q->z32, N, n, st);
    }
    return rc;
}

static int lp_solve_v2_any(const dplrh_plan *p, plr_lp *q, int n, int first, double *X, int ldx,
                           const double *r, int ldr, const double *pm, int pm_n, double *nrm,
                           hipStream_t st);
static bool lp_v2_ok(const dplrh_plan *p, const plr_lp *q);

// r = B - A X (r = B when X is still zero) and partial maxima of |r| per
// column (*pm_n of them, column stride *pm_n).  With `direct` the first
// residual is B itself, read in place, and later ones come out of one fused
// multi-RHS product that reads B for its beta term and also writes the
// maxima of |r| per (column, 128-row block); when that product declines the
// call (an unaligned X, say), or without `direct`, r is B copied, the
// product subtracted, and the maxima taken by k_lp_colmax.
static int lp_residual(dplrh_plan *p, plr_lp *q, int nc, const double *B, int ldb,
                       const double *X, int ldx, bool zero_x, bool direct, const double **r,
                       int *ldr, const double **pm, int *pm_n, hipStream_t st)
{
    const int N = p->N;
    if (direct) {
        if (zero_x) {
            hipLaunchKernelGGL(k_lp_colmax, dim3(LP_SCALE_CHUNKS, nc), dim3(256), 0, st, B, ldb, N,
                               q->pmax);
            PLR_HIP(hipGetLastError());
            *r = B;
            *ldr = ldb;
            *pm = q->pmax;
            *pm_n = LP_SCALE_CHUNKS;
            return 0;
        }
        const int rc = plr_mm_apply_c(p, 0, nc, -1.0, 1.0, X, ldx, B, ldb, q->r, N, p->w,
                                      p->w_per_col, q->bmax, st);
        if (rc < 0) return rc;
        if (rc == 0) {
            *r = q->r;
            *ldr = N;
            *pm = q->bmax;
            *pm_n = N / 128;
            return 0;
        }
    }
    PLR_HIP(hipMemcpy2DAsync(q->r, (size_t)N * sizeof(double), B, (size_t)ldb * sizeof(double),
                             (size_t)N * sizeof(double), nc, hipMemcpyDeviceToDevice, st));
    if (!zero_x) {
        const int rc = plr_gemm_direct(p, 0, nc, -1.0, 1.0, X, ldx, q->r, N, st);
        if (rc) return rc < 0 ? rc : -1;
    }
    hipLaunchKernelGGL(k_lp_colmax, dim3(LP_SCALE_CHUNKS, nc), dim3(256), 0, st, q->r, N, N,
                       q->pmax);
    PLR_HIP(hipGetLastError());
    *r = q->r;
    *ldr = N;
    *pm = q->pmax;
    *pm_n = LP_SCALE_CHUNKS;
    return 0;
}

// X = (first ? 0 : X) + max|r| * solve_lp(r / max|r|), one low-precision
// correction; with nrm, the step's norms (largest |dx| and |x| per column)
// are left in nrm.
static int lp_correct(dplrh_plan *p, plr_lp *q, int nc, int first, double *X, int ldx,
                      const double *r, int ldr, const double *pm, int pm_n, double *nrm,
                      hipStream_t st)
{
    const int N = p->N;
    // (the first kernel of either form zeroes nrm)
    int rc = lp_solve_v2_any(p, q, nc, first, X, ldx, r, ldr, pm, pm_n, nrm, st);
    if (rc <= 0) return rc;
    if (q->type == 3) { // the factorization checked the shape; this is not reached
        plr_set_error("f16 pair storage has no per-level solve");
        return -5;
    }
    // first form: the per-level kernels on r / max|r| in FP32
    hipLaunchKernelGGL(k_lp_scale_rhs, dim3((N + 255) / 256, nc), dim3(256), 0, st, r, ldr, N, pm,
                       pm_n, q->scale, q->r32, nrm);
    PLR_HIP(hipGetLastError());
    rc = q->type == 1 ? lp_solve_t<_Float16>(p, q, nc, st) : lp_solve_t<__bf16>(p, q, nc, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_lp_update, dim3((N + 255) / 256, nc), dim3(256), 0, st, q->z32, N,
                       q->scale, X, ldx, nc, first, nrm);
    PLR_HIP(hipGetLastError());
    return 0;
}

// Error estimate of a refinement step from its correction d_k (k >= 1),
// normwise over the chunk's columns: delta_k = |D_k|_F / |X_k|_F is the
// relative size of the correction, an estimate of the error of X_{k-1}; the
// observed contraction rho_k = delta_k / delta_{k-1} (delta_0 = 1: X_0 is
// the whole first correction) then estimates the relative error of X_k as
// LP_EST_MARGIN * delta_k * min(1, rho_k).  The margin covers the growth of
// the contraction from one step to the next as the error concentrates in
// the directions the low-precision solve contracts least (measured on the
// seeded random family: 9-13% per step).  NaN propagates.
static constexpr double LP_EST_MARGIN = 1.25;

// Reads the sums of steps k0 .. k1 (slots k0 .. k1, one copy; synchronises
// st), carries delta through them and returns the estimate of step k1.
static int lp_estimate(plr_lp *q, int nc, int k0, int k1, double *dprev, double *est,
                       hipStream_t st)
{
    const size_t per = (size_t)LP_NRM_COPIES * 2 * nc;
    PLR_HIP(hipMemcpyAsync(q->nrm_host + (size_t)k0 * per, q->nrm + (size_t)k0 * per,
                           (size_t)(k1 - k0 + 1) * per * sizeof(double), hipMemcpyDeviceToHost, st));
    PLR_HIP(hipStreamSynchronize(st));
    double e = INFINITY;
    for (int k = k0; k <= k1; ++k) {
        const double *h = q->nrm_host + (size_t)k * per;
        double dd = 0.0, xx = 0.0;
        for (size_t i = 0; i < per; i += 2) {
            dd += h[i];
            xx += h[i + 1];
        }
        const double delta = dd == 0.0 ? 0.0 : sqrt(dd) / sqrt(xx); // xx = 0 < dd: inf
        double rho = k == 1 ? delta : (*dprev > 0.0 ? delta / *dprev : (delta == 0.0 ? 0.0 : INFINITY));
        if (rho > 1.0) rho = 1.0;
        *dprev = delta;
        e = LP_EST_MARGIN * delta * rho;
    }
    *est = e;
    return 0;
}

// X := A^{-1} B by mixed-precision solves refined with FP64 residuals
// r = B - A X between them (iterative refinement), column chunks of at most
// 128.  tol < 0: exactly `iters` solves.  tol >= 0: at most `iters`
// (<= LP_NRM_SLOTS) solves, stopping after the first one (from the second
// on) whose error estimate (lp_estimate, normwise over the chunk) is <= tol;
// *used and *err (optional) receive the largest number of solves and
// estimate over the chunks.  The estimates are read back (a synchronisation
// each) from the solve count the previous call on the plan with the same
// tol needed on, the steps before it in one read: a matrix refined again
// with similar right-hand sides synchronises once per call.  res[c] (host, optional) receives
// max|b - A x| of column c after the last solve (one more FP64 residual).
static int lp_solve_body(dplrh_plan *p, plr_lp *q, int n, const double *dB, int ldb, double *dX,
                         int ldx, int iters, double tol, int *used, double *err, double *res,
                         hipStream_t st, const double *dXa = nullptr, int ldxa = 0,
                         double *dYa = nullptr, int ldya = 0)
{
    const int N = p->N;
    if (used) *used = 0;
    if (err) *err = INFINITY;
    if (N == 0 || n == 0) {
        if (err) *err = 0.0;
        return 0;
    }
    const bool adaptive = tol >= 0.0;
    int rc = 0;
    const int chunk = n < 128 ? n : 128;
    if ((rc = lp_ws(p, q, chunk))) return rc;
    const bool direct = lp_v2_ok(p, q) && ldb % 2 == 0 && ((uintptr_t)dB & 15) == 0 && N % 256 == 0;
    if (direct) {
        if ((rc = plr_workspace(p, (chunk + 63) & ~63))) return rc;
        if (!q->bmax) PLR_HIP(hipMalloc(&q->bmax, (size_t)128 * (N / 128) * sizeof(double)));
    }
    int used_max = 0;
    double est_max = adaptive && iters >= 2 ? 0.0 : INFINITY;
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int nc = n - c0 < chunk ? n - c0 : chunk;
        const double *B = dB + (size_t)c0 * ldb;
        double *X = dX + (size_t)c0 * ldx;
        int made = 0, unread = 1; // first step whose sums are not read yet
        double est = INFINITY, dprev = 1.0;
        const size_t per = (size_t)LP_NRM_COPIES * 2 * nc;
        // an apply Ya = A Xa of the same columns rides on the first residual
        // product (one pass over the factors for both) when that product is
        // the fused one; otherwise it is made apart
        const double *Xa = dXa ? dXa + (size_t)c0 * ldxa : nullptr;
        double *Ya = dYa ? dYa + (size_t)c0 * ldya : nullptr;
        bool applied = !dXa;
        for (int it = 0; it < iters; ++it) {
            const double *r, *pm;
            int ldr, pm_n;
            rc = 1;
            if (!applied && it == 1 && direct && nc <= 64) {
                rc = plr_mm_apply_split(p, nc, Xa, ldxa, Ya, ldya, nc, X, ldx, B, ldb, q->r, N,
                                        q->bmax, st);
                if (rc < 0) return rc;
                if (rc == 0) {
                    applied = true;
                    r = q->r;
                    ldr = N;
                    pm = q->bmax + (size_t)64 * (N / 128);
                    pm_n = N / 128;
                }
            }
            if (rc == 1 &&
                (rc = lp_residual(p, q, nc, B, ldb, X, ldx, it == 0, direct, &r, &ldr, &pm, &pm_n, st)))
                return rc;
            const bool track = adaptive && it >= 1;
            rc = lp_correct(p, q, nc, it == 0, X, ldx, r, ldr, pm, pm_n,
                            track ? q->nrm + (size_t)it * per : nullptr, st);
            if (rc) return rc < 0 ? rc : -1;
            made = it + 1;
            if (!track || (tol == q->pred_tol && made < q->pred && it + 1 < iters)) continue;
            if ((rc = lp_estimate(q, nc, unread, it, &dprev, &est, st))) return rc;
            unread = it + 1;
            if (est <= tol) break;
        }
        if (!applied && (rc = plr_gemm_direct(p, 0, nc, 1.0, 0.0, Xa, ldxa, Ya, ldya, st)))
            return rc < 0 ? rc : -1;
        if (made > used_max) used_max = made;
        if (adaptive && (est > est_max || est != est)) est_max = est;
        if (res) {
            const double *r, *pm;
            int ldr, pm_n;
            if ((rc = lp_residual(p, q, nc, B, ldb, X, ldx, false, direct, &r, &ldr, &pm, &pm_n, st)))
                return rc;
            hipLaunchKernelGGL(k_lp_pmfold, dim3((nc + 63) / 64), dim3(64), 0, st, pm, pm_n, nc,
                               q->scale);
            PLR_HIP(hipGetLastError());
            PLR_HIP(hipMemcpyAsync(res + c0, q->scale, nc * sizeof(double), hipMemcpyDeviceToHost, st));
            PLR_HIP(hipStreamSynchronize(st));
        }
    }
    if (adaptive) {
        q->pred = used_max;
        q->pred_tol = tol;
    }
    if (used) *used = used_max;
    if (err) *err = adaptive ? est_max : INFINITY;
    return 0;
}

// Launched directly: with kernels of 5-100 us the stream stays fed, and
// measured, the graph's own start cost more than it saved here.
extern "C" int dplrh_lp_solve(dplrh_plan *p, int n, const double *dB, int ldb, double *dX,
                              int ldx, int iters, double *res, void *stream)
{
    plr_lp *q = lp_of(p);
    if (!q) {
        plr_set_error("solve before a successful mixed-precision factorization");
        return -4;
    }
    if (iters < 1) {
        plr_set_error("refined solve needs at least one iteration (got %d)", iters);
        return -5;
    }
    return lp_solve_body(p, q, n, dB, ldb, dX, ldx, iters, -1.0, nullptr, nullptr, res,
                         (hipStream_t)stream);
}

// The apply Ya = A Xa and the refined solve X = A^{-1} B of n columns in one
// call (the mixed-precision step): the apply shares the factor reads of the
// first residual product.
extern "C" int dplrh_lp_apply_solve_tol(dplrh_plan *p, int n, const double *dXa, int ldxa,
                                        double *dYa, int ldya, const double *dB, int ldb,
                                        double *dX, int ldx, double tol, int maxit, int *iters,
                                        double *err, void *stream)
{
    plr_lp *q = lp_of(p);
    if (!q) {
        plr_set_error("solve before a successful mixed-precision factorization");
        return -4;
    }
    if (!(tol >= 0.0) || maxit < 2 || maxit > LP_NRM_SLOTS) {
        plr_set_error("refined solve: tol must be >= 0 and 2 <= maxit <= %d (got %g, %d)",
                      LP_NRM_SLOTS, tol, maxit);
        return -5;
    }
    if (!dXa || !dYa || ldxa < p->N || ldya < p->N) {
        plr_set_error("apply and solve: the apply's operands are missing or too short");
        return -5;
    }
    return lp_solve_body(p, q, n, dB, ldb, dX, ldx, maxit, tol, iters, err, nullptr,
                         (hipStream_t)stream, dXa, ldxa, dYa, ldya);
}

extern "C" int dplrh_lp_solve_tol(dplrh_plan *p, int n, const double *dB, int ldb, double *dX,
                                  int ldx, double tol, int maxit, int *iters, double *err,
                                  void *stream)
{
    plr_lp *q = lp_of(p);
    if (!q) {
        plr_set_error("solve before a successful mixed-precision factorization");
        return -4;
    }
    if (!(tol >= 0.0) || maxit < 2 || maxit > LP_NRM_SLOTS) {
        plr_set_error("refined solve: tol must be >= 0 and 2 <= maxit <= %d (got %g, %d)",
                      LP_NRM_SLOTS, tol, maxit);
        return -5;
    }
    return lp_solve_body(p, q, n, dB, ldb, dX, ldx, maxit, tol, iters, err, nullptr,
                         (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Solve, second form: the bottom levels in LDS, then one pass per upper level.
//
//   k_lp_tail  one work-group per (span of S rows, CW right-hand sides):
//              z = r / max|r| per column, z|_leaf = Dinv z|_leaf, then every
//              level d >= d0 (segments inside the span) with z in LDS, and
//              finally T of level d0 - 1 (the span is one of its segments).
//   k_lp_level one work-group per (LV_RB = 256 rows, 64 right-hand sides)
//              of level d < d0: the rows of inv(K) for its segment
//              (recomputed per work-group), z += Ut_d W_d, then the partial
//              sums of V_{d-1}^T z over its rows (summed by k_lp_finish1), or,
//              at d = 0, x (+)= max|r| * z in FP64.
// Two kernels per upper level instead of four, and one pass over z.

static constexpr int LT_THREADS = 1024;          // sixteen waves
static constexpr int LT_NW = LT_THREADS / 64;
static constexpr int LT_MAXT = 512 / 16 / LT_NW; // 16-row leaf tiles per wave (span <= 512)
static constexpr int LT_ZU = 16 * 512 / LT_THREADS; // 16-byte r loads per thread and round
static constexpr int LV_RB = 256, LV_CW = 64, LV_THREADS = 512;
static constexpr int LV_RT = LV_RB / 128; // 16-row tiles per wave

struct LpTailTab
{
    int N, b, L, d0, span, RT, tsegs; // RT: largest Rp over levels d0-1 .. L-1
    int Rp[PLR_MAXL];
    const void *V[PLR_MAXL], *Ut[PLR_MAXL];
    const float *S[PLR_MAXL];
    const float *K32[PLR_MAXL];
    const double *K64[PLR_MAXL];
    const void *D;
    // pair storage: lo planes at these element offsets from V, Ut, D, and the
    // inverse factor scales (1 for single storage)
    size_t vlo, ulo, dlo;
    float isV[PLR_MAXL], isU[PLR_MAXL], isD;
    // reassociated upper levels: V_d^T z over the span for every level
    // d < d0, at upart + upoff[d] + (span * Rp[d] + k) * ncols + c
    float *upart;
    size_t upoff[PLR_MAXL];
    int nspan;
};

struct LpLvlTab
{
    int N, R, Rn, seg, segn, tps, nkg, ncg;
    const void *Ut, *Vn;
    const float *S;
    const float *K32;
    const double *K64;
    size_t vlo, ulo;   // pair storage: lo planes of Vn and Ut
    float isU, isVn;   // inverse factor scales
};


// smallest power of two >= v (1 for v = 0); v >= 0 as the bits of a float
__device__ __forceinline__ float lp_pow2(unsigned bits)
{
    const float v = __uint_as_float(bits);
    if (!(v > 0.f)) return v == 0.f ? 1.f : v; // NaN stays NaN
    int e;
    const float f = frexpf(v, &e);
    return ldexpf(1.f, f == 0.5f ? e - 1 : e);
}

// Capacitance rows for P tasks at once: task p computes the 16 x 16 tile
// c = inv(K)[i0[p] .. +15, :] rhs[:, 16 columns], where row j of rhs is
// lo[p][j * rs + m] for j < R and hi[p][(j - R) * rs + m] above (LDS), and
// multiplies row i by its sig (sv[p] points at the sig of row i0[p] + g;
// row + 4v is sv[p][4v]).  Every K and sig load of the P tasks is issued
// before the first MFMA.  K is column-major n2 x n2, stored in FP64 (K64)
// or FP32 (K32); the arithmetic is FP64 MFMA either way (an FP32 product
// loses digits to cancellation in inv(K) rhs that the refinement then has
// to win back).  Lane result, in the FP64 MFMA's D layout:
// w[p][v] = -sig * c[i0[p] + g + 4v][m].
template <int P>
__device__ __forceinline__ void lp_cap_tiles(const float *const *K32, const double *const *K64,
                                             int n2, int R, const int *i0,
                                             const float *const *lo, const float *const *hi,
                                             int rs, const float *const *sv, int m, int g,
                                             f4 *w)
{
    float s[P][4];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) s[p][v] = sv[p][4 * v];
    plr_d4 acc[P];
#pragma unroll
    for (int p = 0; p < P; ++p) acc[p] = plr_d4{0.0, 0.0, 0.0, 0.0};
    if (K64[0]) {
        for (int j0 = 0; j0 < n2; j0 += 64) {
            double av[P][16];
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    av[p][u] = K64[p][(size_t)(j0 + 4 * u + g) * n2 + i0[p] + m];
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int j = j0 + 4 * u + g;
                    const float *row = j < R ? lo[p] + j * rs : hi[p] + (j - R) * rs;
                    acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[p][u], (double)row[m], acc[p], 0, 0, 0);
                }
        }
    } else {
        for (int j0 = 0; j0 < n2; j0 += 64) {
            float av[P][16];
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    av[p][u] = K32[p][(size_t)(j0 + 4 * u + g) * n2 + i0[p] + m];
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int j = j0 + 4 * u + g;
                    const float *row = j < R ? lo[p] + j * rs : hi[p] + (j - R) * rs;
                    acc[p] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[p][u], (double)row[m],
                                                                  acc[p], 0, 0, 0);
                }
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) w[p][v] = (float)(-(double)s[p][v] * acc[p][v]);
}

// Eight consecutive floats of LDS as two 16-byte reads (rows are padded to
// 16 bytes, not 32).
__device__ __forceinline__ f8 lds8(const float *p)
{
    const f4 lo = *(const f4 *)p, hi = *(const f4 *)(p + 4);
    return f8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

// LDS row strides (floats) are chosen for the access patterns: rows read
// by 16 lanes at once with 16-byte accesses have an odd number of 16-byte
// units (S + 4, R + 4); rows read one float per lane by four lane groups
// are 16 mod 64 (CW + 16, 64 + 16), so no two lanes of a group share a bank.
// LDS of k_lp_tail: z [CW][S+4], T [tsegs][RT][CW+16], W [tsegs][CW][RT+4],
// smax [tsegs][CW] (floats).
static size_t lpt_lds(int S, int CW, int RT, int tsegs)
{
    return ((size_t)CW * (S + 4) + (size_t)tsegs * RT * (CW + 16) + (size_t)tsegs * CW * (RT + 4) +
            (size_t)tsegs * CW) * sizeof(float);
}

// T[c][k] sums over segments of `seg` rows of the span: sum_i V[k][row0 + i]
// z[c][i], in batches of eight 32-row MFMA steps per wave (the eight V loads
// of a batch go out together).  Segments of at most 256 rows end inside a
// batch and are written by emit(s, kt, ct, acc); longer ones leave one
// partial per batch, emitted with s = -1 - (batch index in the segment).
// With PR (pair storage) V's lo plane is at V + vlo, z is split into hi and
// lo times LP_PAIR_Z, and emitted sums are scaled by os (the inverse of V's
// scale times LP_PAIR_Z).
template <typename T16, int NCT, bool PR, typename Emit>
__device__ __forceinline__ void lp_tail_reduce(const T16 *V, int N, int row0, const float *z,
                                               int zs, int S, int seg, int nkt, int wv, int m,
                                               int g, size_t vlo, float os, Emit emit)
{
    typedef typename Lp<T16>::v8 v8;
    constexpr int QB = PR ? 4 : 8; // loads in flight: pairs take two halves of the batch
    const int nstep = S / 32, spq = seg / 32, total = nstep * nkt * NCT;
    for (int e0 = wv * 8; e0 < total; e0 += 8 * LT_NW) {
      f4 acc = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q0 = 0; q0 < 8; q0 += QB) {
        v8 av[QB], al[QB];
#pragma unroll
        for (int qq = 0; qq < QB; ++qq) {
            const int e = e0 + q0 + qq, grp = e / nstep, rs = e - grp * nstep, kt = grp / NCT;
            const size_t idx = (size_t)(kt * 16 + m) * N + row0 + rs * 32 + 8 * g;
            av[qq] = *(const v8 *)(V + idx);
            if (PR) al[qq] = *(const v8 *)(V + vlo + idx);
        }
#pragma unroll
        for (int qq = 0; qq < QB; ++qq) {
            const int q = q0 + qq;
            const int e = e0 + q, grp = e / nstep, rs = e - grp * nstep;
            const int kt = grp / NCT, ct = grp - kt * NCT;
            const f8 zv = lds8(z + (ct * 16 + m) * zs + rs * 32 + 8 * g);
            if (PR) {
                v8 zh, zl;
                split8<T16>(zv, LP_PAIR_Z, zh, zl);
                acc = mfma3<T16>(av[qq], al[qq], zh, zl, acc);
            } else {
                acc = Lp<T16>::mfma(av[qq], cvt8<T16>(zv, 1.f), acc);
            }
            if ((rs + 1) % spq == 0 && spq <= 8) {
                emit(rs / spq, kt, ct, PR ? acc * os : acc);
                acc = f4{0.f, 0.f, 0.f, 0.f};
            } else if (q == 7 && spq > 8) {
                emit(-1 - (rs % spq) / 8, kt, ct, PR ? acc * os : acc);
            }
        }
      }
    }
}

// pmax: pm_n partial maxima of |r| per column (column stride pm_n).
template <typename T16, int CW, bool PR>
__global__ __launch_bounds__(LT_THREADS) void k_lp_tail(LpTailTab a, const double *__restrict__ r,
                                                        int ldr, const double *__restrict__ pmax,
                                                        int pm_n, double *__restrict__ scale,
                                                        float *__restrict__ Z, int ldz,
                                                        float *__restrict__ Tout, size_t tstride,
                                                        int ncols, double *__restrict__ nrm)
{
    typedef typename Lp<T16>::v8 v8;
    constexpr int NCT = CW / 16;
    extern __shared__ float lds[];
    __shared__ double cinv[CW];
    const int S = a.span, zs = S + 4, RT = a.RT, nsm = a.tsegs;
    const int ts = CW + 16, tsz = RT * ts;  // T[s][k][c]
    const int wsk = RT + 4, wsz = CW * wsk; // W[s][c][k]
    float *z = lds;
    float *T = z + CW * zs;
    float *W = T + nsm * tsz;
    unsigned *smax = (unsigned *)(W + nsm * wsz); // [s][c]
    const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wv = tid >> 6;
    const int row0 = blockIdx.x * S, c0 = blockIdx.y * CW, N = a.N, b = a.b;
    const int nc = min(CW, ncols - c0);

    // 0. the correction norms of this solve start from zero (its last
    // kernel adds into them)
    if (nrm && blockIdx.x == 0)
        for (int e = tid; e < LP_NRM_COPIES * 2 * nc; e += LT_THREADS) {
            const int cp = e / (2 * nc), rem = e - cp * 2 * nc;
            nrm[((size_t)cp * ncols + c0 + rem / 2) * 2 + (rem & 1)] = 0.0;
        }
    // 1. column scales (max |r| from the partial maxima, eight threads per
    // column), then z = r / max|r|
    {
        double *cp = (double *)T; // [CW][8], T is free until the first level
        const int c = tid >> 3, part = tid & 7;
        if (c < CW) {
            const double *pm = pmax + (size_t)(c0 + min(c, nc - 1)) * pm_n;
            double mv = 0.0;
            for (int j0 = part * 8; j0 < pm_n; j0 += 64) {
                double pv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) pv[u] = pm[min(j0 + u, pm_n - 1)];
#pragma unroll
                for (int u = 0; u < 8; ++u) mv = fmax(mv, pv[u]);
            }
            cp[c * 8 + part] = mv;
        }
        __syncthreads();
        if (tid < CW) {
            double mv = 0.0;
#pragma unroll
            for (int j = 0; j < 8; ++j) mv = fmax(mv, cp[tid * 8 + j]);
            if (tid < nc && blockIdx.x == 0) scale[c0 + tid] = mv;
            cinv[tid] = tid < nc ? 1.0 / (mv > 0.0 ? mv : 1.0) : 0.0;
        }
        __syncthreads();
    }
    {
        const int hs = S / 2, tot = CW * hs;
        for (int e0 = 0; e0 < tot; e0 += LT_ZU * LT_THREADS) {
            double2 v[LT_ZU];
#pragma unroll
            for (int u = 0; u < LT_ZU; ++u) {
                const int e = min(e0 + u * LT_THREADS + tid, tot - 1);
                const int c = e / hs, i = 2 * (e - c * hs);
                v[u] = *(const double2 *)(r + (size_t)(c0 + min(c, nc - 1)) * ldr + row0 + i);
            }
#pragma unroll
            for (int u = 0; u < LT_ZU; ++u) {
                const int e = e0 + u * LT_THREADS + tid;
                if (e < tot) {
                    const int c = e / hs, i = 2 * (e - c * hs);
                    z[c * zs + i] = (float)(v[u].x * cinv[c]);
                    z[c * zs + i + 1] = (float)(v[u].y * cinv[c]);
                }
            }
        }
    }
    __syncthreads();

    // 2. leaves: each wave takes up to four 16-row tiles (M = rows, N =
    // columns, K = leaf columns), all D loads of a 64-column step together;
    // results stay in registers until every wave has read its leaf
    {
        const T16 *Dh = (const T16 *)a.D;
        const int ntile = S / 16;
        f4 acc[LT_MAXT][NCT];
        int i0[LT_MAXT], lf[LT_MAXT];
#pragma unroll
        for (int q = 0; q < LT_MAXT; ++q) {
            i0[q] = min(wv + LT_NW * q, ntile - 1) * 16;
            lf[q] = i0[q] / b;
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct) acc[q][ct] = f4{0.f, 0.f, 0.f, 0.f};
        }
        for (int j0 = 0; j0 < b; j0 += 64) {
            v8 a0[LT_MAXT], a1[LT_MAXT], l0[LT_MAXT], l1[LT_MAXT];
#pragma unroll
            for (int q = 0; q < LT_MAXT; ++q) {
                const T16 *D = Dh + ((size_t)(row0 / b + lf[q]) * b + i0[q] - lf[q] * b + m) * b + 8 * g + j0;
                a0[q] = *(const v8 *)D;
                a1[q] = *(const v8 *)(D + 32);
                if (PR) {
                    l0[q] = *(const v8 *)(D + a.dlo);
                    l1[q] = *(const v8 *)(D + a.dlo + 32);
                }
            }
#pragma unroll
            for (int q = 0; q < LT_MAXT; ++q) {
                const float *zb = z + m * zs + lf[q] * b + 8 * g + j0;
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) {
                    const f8 x0 = lds8(zb + ct * 16 * zs);
                    const f8 x1 = lds8(zb + ct * 16 * zs + 32);
                    if (PR) {
                        v8 h0, t0, h1, t1;
                        split8<T16>(x0, LP_PAIR_Z, h0, t0);
                        split8<T16>(x1, LP_PAIR_Z, h1, t1);
                        acc[q][ct] = mfma3<T16>(a0[q], l0[q], h0, t0, acc[q][ct]);
                        acc[q][ct] = mfma3<T16>(a1[q], l1[q], h1, t1, acc[q][ct]);
                    } else {
                        acc[q][ct] = Lp<T16>::mfma(a0[q], cvt8<T16>(x0, 1.f), acc[q][ct]);
                        acc[q][ct] = Lp<T16>::mfma(a1[q], cvt8<T16>(x1, 1.f), acc[q][ct]);
                    }
                }
            }
        }
        if (PR) {
            const float os = a.isD / LP_PAIR_Z;
#pragma unroll
            for (int q = 0; q < LT_MAXT; ++q)
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) acc[q][ct] *= os;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LT_MAXT; ++q)
            if (wv + LT_NW * q < ntile)
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct)
                    *(f4 *)(z + (ct * 16 + m) * zs + i0[q] + 4 * g) = acc[q][ct];
        __syncthreads();
    }

    // 3. levels L-1 .. d0
    for (int d = a.L - 1; d >= a.d0; --d) {
        const int R = a.Rp[d], seg = N >> (d + 1), nsg = S / seg, nkt = R / 16;
        if (R == 0) continue;
        const int sg0 = row0 / seg; // global index of the span's first segment
        const T16 *V = (const T16 *)a.V[d], *U = (const T16 *)a.Ut[d];
        // (a) T[s][k][c] = sum over segment s of V[k][i] z[c][i]
        for (int e = tid; e < nsg * CW; e += LT_THREADS) smax[e] = 0u;
        lp_tail_reduce<T16, NCT, PR>(V, N, row0, z, zs, S, seg, nkt, wv, m, g, a.vlo,
                                     a.isV[d] / LP_PAIR_Z, [&](int s, int kt, int ct, f4 acc) {
                                     float *t = T + s * tsz + (kt * 16 + 4 * g) * ts + ct * 16 + m;
#pragma unroll
                                     for (int v = 0; v < 4; ++v) t[v * ts] = acc[v];
                                 });
        __syncthreads();
        // (b) c = inv(K) [T_bot; T_top] per node, W_top = -S_top .* c[0:R],
        // W_bot = -S_bot .* c[R:2R], with the largest |W| per (segment, column);
        // two tasks per wave and pass
        {
            const int n2 = 2 * R, nit = n2 / 16, ntask = (nsg / 2) * nit * NCT;
            for (int t0 = wv * 2; t0 < ntask; t0 += 2 * LT_NW) {
                const float *k32[2], *lo[2], *hi[2], *sv[2];
                const double *k64[2];
                int i0[2], sl[2], ct[2], kk[2];
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    const int task = min(t0 + p, ntask - 1);
                    ct[p] = task % NCT;
                    const int it = (task / NCT) % nit, nl = task / (NCT * nit);
                    const int top = 2 * nl, bot = top + 1;
                    const size_t ko = (size_t)(sg0 / 2 + nl) * n2 * n2;
                    k32[p] = a.K32[d] ? a.K32[d] + ko : nullptr;
                    k64[p] = a.K64[d] ? a.K64[d] + ko : nullptr;
                    i0[p] = it * 16;
                    lo[p] = T + bot * tsz + ct[p] * 16;
                    hi[p] = T + top * tsz + ct[p] * 16;
                    // a lane's rows it*16 + g + 4v lie in one half (R is a
                    // multiple of 16)
                    const int i = it * 16 + g, h = it * 16 >= R;
                    kk[p] = h ? i - R : i;
                    sl[p] = h ? bot : top;
                    sv[p] = a.S[d] + (size_t)(sg0 + sl[p]) * R + kk[p];
                }
                f4 w[2];
                lp_cap_tiles<2>(k32, k64, n2, R, i0, lo, hi, ts, sv, m, g, w);
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    if (t0 + p >= ntask) continue;
                    float *wr = W + sl[p] * wsz + (ct[p] * 16 + m) * wsk + kk[p];
#pragma unroll
                    for (int v = 0; v < 4; ++v) wr[4 * v] = w[p][v];
                    const float mx = fmaxf(fmaxf(fabsf(w[p][0]), fabsf(w[p][1])),
                                           fmaxf(fabsf(w[p][2]), fabsf(w[p][3])));
                    atomicMax(&smax[sl[p] * CW + ct[p] * 16 + m], __float_as_uint(mx));
                }
            }
        }
        __syncthreads();
        // (c) z[c][i] += sum_k Ut[i][k] W[s(i)][c][k]  (M = columns, N = rows),
        // batches of eight MFMA steps per wave with their Ut loads together
        {
            const int nks = R / 32, total = (S / 16) * NCT * nks;
            constexpr int QB = PR ? 4 : 8; // loads in flight (pairs: two halves of the batch)
            const float os = PR ? a.isU[d] / LP_PAIR_Z : 1.f;
            for (int e0 = wv * 8; e0 < total; e0 += 8 * LT_NW) {
              f4 acc = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
              for (int q0 = 0; q0 < 8; q0 += QB) {
                v8 bq[QB], bl[QB];
#pragma unroll
                for (int qq = 0; qq < QB; ++qq) {
                    const int e = e0 + q0 + qq, ks = e % nks, it = (e / nks) / NCT;
                    const size_t idx = (size_t)(row0 + it * 16 + m) * R + ks * 32 + 8 * g;
                    bq[qq] = *(const v8 *)(U + idx);
                    if (PR) bl[qq] = *(const v8 *)(U + a.ulo + idx);
                }
#pragma unroll
                for (int qq = 0; qq < QB; ++qq) {
                    const int q = q0 + qq;
                    const int e = e0 + q, ks = e % nks, grp = e / nks, it = grp / NCT, ct = grp - it * NCT;
                    const int s = (it * 16) / seg;
                    const unsigned *sm = smax + s * CW + ct * 16;
                    const f8 w8 = lds8(W + s * wsz + (ct * 16 + m) * wsk + ks * 32 + 8 * g);
                    if (PR) {
                        v8 wh, wl;
                        split8<T16>(w8, LP_PAIR_Z / lp_pow2(sm[m]), wh, wl);
                        acc = mfma3<T16>(wh, wl, bq[qq], bl[qq], acc);
                    } else {
                        acc = Lp<T16>::mfma(cvt8<T16>(w8, 1.f / lp_pow2(sm[m])), bq[qq], acc);
                    }
                    if (ks == nks - 1) {
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            z[(ct * 16 + 4 * g + v) * zs + it * 16 + m] += acc[v] * (os * lp_pow2(sm[4 * g + v]));
                        acc = f4{0.f, 0.f, 0.f, 0.f};
                    }
                }
              }
            }
        }
        __syncthreads();
    }

    // 4. T of level d0 - 1 over the span (one of its segments), or with
    // upart the span's share of T of level d0 - 1 and of every coarser
    // level; a span of more than 256 rows leaves partials combined in a
    // fixed order
    for (int dn = a.d0 - 1; dn >= (a.upart ? 0 : a.d0 - 1); --dn) {
        const int R = a.Rp[dn], nkt = R / 16, sg = row0 / S;
        const int nh = S > 256 ? S / 256 : 1; // partials per (kt, ct)
        float *red = T;                        // free now: [kt][ct][h][lane][4]
        float *up = a.upart ? a.upart + a.upoff[dn] + (size_t)sg * R * ncols : nullptr;
        // rows k .. k+3 of column c
        auto put = [&](int c, int k, f4 x) {
            if (up) {
#pragma unroll
                for (int v = 0; v < 4; ++v) up[(size_t)(k + v) * ncols + c0 + c] = x[v];
            } else {
                *(f4 *)(Tout + (size_t)(c0 + c) * tstride + (size_t)sg * R + k) = x;
            }
        };
        if (R > 0) {
            lp_tail_reduce<T16, NCT, PR>((const T16 *)a.V[dn], N, row0, z, zs, S, S, nkt, wv, m, g,
                                         a.vlo, a.isV[dn] / LP_PAIR_Z,
                                         [&](int s, int kt, int ct, f4 acc) {
                                         if (s >= 0) {
                                             const int c = ct * 16 + m;
                                             if (c < nc) put(c, kt * 16 + 4 * g, acc);
                                         } else {
                                             const int h = -1 - s;
                                             *(f4 *)(red + ((((kt * NCT + ct) * nh + h) * 64) + lane) * 4) = acc;
                                         }
                                     });
        }
        if (nh > 1) {
            __syncthreads();
            for (int tt = wv; tt < nkt * NCT; tt += LT_NW) {
                const int kt = tt / NCT, ct = tt - kt * NCT, c = ct * 16 + m;
                f4 x = *(const f4 *)(red + ((tt * nh) * 64 + lane) * 4);
                for (int h = 1; h < nh; ++h) x += *(const f4 *)(red + (((tt * nh + h) * 64) + lane) * 4);
                if (c < nc) put(c, kt * 16 + 4 * g, x);
            }
            __syncthreads(); // red is reused by the next level
        }
    }

    // 5. z of the span
    for (int e = tid; e < CW * (S / 4); e += LT_THREADS) {
        const int c = e / (S / 4), i = 4 * (e - c * (S / 4));
        if (c < nc) *(f4 *)(Z + (size_t)(c0 + c) * ldz + row0 + i) = *(const f4 *)(z + c * zs + i);
    }
}

// LDS of k_lp_level: rhs [2R][CW+16], W [CW][R+4] (floats), zt [CW][RB+8]
// (16-bit; with pair storage a second zt for the lo parts)
static size_t lpl_lds(int R, bool pair)
{
    return ((size_t)2 * R * (LV_CW + 16) + (size_t)LV_CW * (R + 4)) * sizeof(float) +
           (size_t)LV_CW * (LV_RB + 8) * 2 * (pair ? 2 : 1);
}

// Every global load of the work-group (old z, Ut, V of level d-1, T, K, sig)
// is issued at the start, so the work-group waits on memory about once.  Eight
// waves of LV_RT 16-row tiles each; RM bounds the ranks of levels d and d-1
// (32 or 64), which sizes the register buffers.
template <typename T16, int RM, bool PR>
__global__ __launch_bounds__(LV_THREADS) void k_lp_level(LpLvlTab a, const float *__restrict__ Tin,
                                                         size_t tstride, float *__restrict__ Z,
                                                         int ldz, float *__restrict__ part,
                                                         const double *__restrict__ scale,
                                                         double *__restrict__ X, int ldx, int first,
                                                         int ncols, double *__restrict__ nrm)
{
    typedef typename Lp<T16>::v8 v8;
    constexpr int KS = RM / 32;    // expansion k-steps
    constexpr int NP = RM / 32;    // capacitance and reduction tasks per wave
    constexpr int NU = LV_RB / 32; // reduction k-steps (rows)
    extern __shared__ float lds[];
    __shared__ unsigned smax[LV_CW];
    __shared__ double xnrm[2 * LV_CW]; // X pass: sums of dx^2, x^2 per column
    const int R = a.R, n2 = 2 * R, rs = LV_CW + 16, wsk = R + 4, zts = LV_RB + 8;
    float *rhs = lds;                   // [n2][rs]
    float *W = rhs + n2 * rs;           // [CW][wsk]
    T16 *zt = (T16 *)(W + LV_CW * wsk); // [CW][zts]
    T16 *ztl = zt + LV_CW * zts;        // [CW][zts], lo parts (pair storage)
    const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wv = tid >> 6;
    const int r0 = blockIdx.x * LV_RB, c0 = blockIdx.y * LV_CW, N = a.N;
    const int ri = r0 + wv * 16 * LV_RT; // this wave's rows
    const int nc = min(LV_CW, ncols - c0);
    const int s = r0 / a.seg, istop = !(s & 1), top = s & ~1, bot = top + 1;
    const int nks = R / 32, nkt = R / 16, nktn = a.Rn / 16;

    // old z as the expansion's D (M = rows, N = columns): rows ri + 16rt +
    // 4g + v of column 16ct + m, one 16-byte load per tile (ldz % 4 == 0)
    f4 zo[LV_RT][4];
#pragma unroll
    for (int rt = 0; rt < LV_RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
            zo[rt][ct] = *(const f4 *)(Z + (size_t)(c0 + min(ct * 16 + m, nc - 1)) * ldz + ri + 16 * rt + 4 * g);
    // Ut rows of the expansion
    const T16 *U = (const T16 *)a.Ut;
    v8 bq[LV_RT][KS], bl[LV_RT][KS];
    if (R > 0) // a level of rank 0 has no Ut rows
#pragma unroll
        for (int rt = 0; rt < LV_RT; ++rt)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const size_t idx = (size_t)(ri + 16 * rt + m) * R + min(ks, nks - 1) * 32 + 8 * g;
                bq[rt][ks] = *(const v8 *)(U + idx);
                if (PR) bl[rt][ks] = *(const v8 *)(U + a.ulo + idx);
            }
    // V of level d-1 for this wave's reduction tasks (kt, ct) = wv + 8p
    v8 av[NP][NU], avl[NP][NU];
    if (a.Rn > 0) {
        const T16 *Vn = (const T16 *)a.Vn;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int task = min(wv + 8 * p, nktn * 4 - 1), kt = task >> 2;
#pragma unroll
            for (int u = 0; u < NU; ++u) {
                const size_t idx = (size_t)(kt * 16 + m) * N + r0 + 32 * u + 8 * g;
                av[p][u] = *(const v8 *)(Vn + idx);
                if (PR) avl[p][u] = *(const v8 *)(Vn + a.vlo + idx);
            }
        }
    }
    // 1. rhs = [T_bot; T_top] of the node
    for (int e0 = 0; e0 < n2 * LV_CW; e0 += 8 * LV_THREADS) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = min(e0 + u * LV_THREADS + tid, n2 * LV_CW - 1);
            const int cl = e / n2, j = e - cl * n2, hi = j >= R;
            const int sg = hi ? top : bot, kk = hi ? j - R : j;
            v[u] = Tin[(size_t)(c0 + min(cl, nc - 1)) * tstride + (size_t)sg * R + kk] *
                   (cl < nc ? 1.f : 0.f);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * LV_THREADS + tid;
            if (e < n2 * LV_CW) {
                const int cl = e / n2, j = e - cl * n2;
                rhs[j * rs + cl] = v[u];
            }
        }
    }
    if (tid < LV_CW) smax[tid] = 0u;
    if (tid < 2 * LV_CW) xnrm[tid] = 0.0;
    __syncthreads();
    // 2. W_s = -S_s .* (rows of inv(K) for segment s) rhs, largest |W| per
    // column: tasks (kt, ct), wave wv takes wv + 8p
    if (R > 0) {
        const size_t ko = (size_t)(s >> 1) * n2 * n2;
        const int ntask = nkt * 4;
        const float *k32[NP], *lo[NP], *hi[NP], *sv[NP];
        const double *k64[NP];
        int i0[NP], ct[NP], kk[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int task = min(wv + 8 * p, ntask - 1);
            ct[p] = task & 3;
            kk[p] = (task >> 2) * 16;
            i0[p] = (istop ? 0 : R) + kk[p];
            k32[p] = a.K64 ? nullptr : a.K32 + ko;
            k64[p] = a.K64 ? a.K64 + ko : nullptr;
            lo[p] = rhs + ct[p] * 16;
            hi[p] = rhs + R * rs + ct[p] * 16;
            kk[p] += g; // rows kk + 4v (the FP64 MFMA's D layout)
            sv[p] = a.S + (size_t)s * R + kk[p];
        }
        f4 w[NP];
        lp_cap_tiles<NP>(k32, k64, n2, R, i0, lo, hi, rs, sv, m, g, w);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            if (wv + 8 * p >= ntask) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) W[(ct[p] * 16 + m) * wsk + kk[p] + 4 * v] = w[p][v];
            const float mx = fmaxf(fmaxf(fabsf(w[p][0]), fabsf(w[p][1])), fmaxf(fabsf(w[p][2]), fabsf(w[p][3])));
            atomicMax(&smax[ct[p] * 16 + m], __float_as_uint(mx));
        }
    }
    __syncthreads();
    // 3. z += Ut W (M = rows, N = columns): LV_RT 16-row tiles per wave
    {
        const bool xal = X && ((uintptr_t)X & 15) == 0 && ldx % 2 == 0;
        float isc[4], sc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            sc[ct] = lp_pow2(smax[ct * 16 + m]);
            isc[ct] = 1.f / sc[ct];
        }
        // pair storage: W split times LP_PAIR_Z; the sums come back times Ut's
        // inverse scale over LP_PAIR_Z (column scales sco, one per lane's
        // four columns in the D layout below)
        float sco[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) sco[ct] = PR ? sc[ct] * (a.isU / LP_PAIR_Z) : sc[ct];
        f4 acc[LV_RT][4];
#pragma unroll
        for (int rt = 0; rt < LV_RT; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            if (ks < nks) {
#pragma unroll
                for (int ct = 0; ct < 4; ++ct) {
                    const f8 w8 = lds8(W + (ct * 16 + m) * wsk + ks * 32 + 8 * g);
                    if (PR) {
                        v8 wh, wl;
                        split8<T16>(w8, isc[ct] * LP_PAIR_Z, wh, wl);
#pragma unroll
                        for (int rt = 0; rt < LV_RT; ++rt)
                            acc[rt][ct] = mfma3<T16>(bq[rt][ks], bl[rt][ks], wh, wl, acc[rt][ct]);
                    } else {
                        const v8 bw = cvt8<T16>(w8, isc[ct]);
#pragma unroll
                        for (int rt = 0; rt < LV_RT; ++rt) acc[rt][ct] = Lp<T16>::mfma(bq[rt][ks], bw, acc[rt][ct]);
                    }
                }
            }
        }
        // lane holds dz[row i = 16 (LV_RT wv + rt) + 4g + v][column c = 16ct + m]
        double dmx[4] = {0.0, 0.0, 0.0, 0.0}, xmx[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int rt = 0; rt < LV_RT; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int c = ct * 16 + m, i = (wv * LV_RT + rt) * 16 + 4 * g;
                const float on = c < nc ? 1.f : 0.f;
                f4 zn;
#pragma unroll
                for (int v = 0; v < 4; ++v) zn[v] = zo[rt][ct][v] * on + acc[rt][ct][v] * sco[ct];
                if (X) {
                    if (c < nc) {
                        double *xp = X + (size_t)(c0 + c) * ldx + r0 + i;
                        const double sr = scale[c0 + c] > 0.0 ? scale[c0 + c] : 1.0;
                        if (xal) { // four consecutive rows as two 16-byte accesses
#pragma unroll
                            for (int h = 0; h < 2; ++h) {
                                double2 o = first ? double2{0.0, 0.0} : *(const double2 *)(xp + 2 * h);
                                const double d0 = sr * (double)zn[2 * h], d1 = sr * (double)zn[2 * h + 1];
                                o.x += d0;
                                o.y += d1;
                                *(double2 *)(xp + 2 * h) = o;
                                dmx[ct] += d0 * d0 + d1 * d1;
                                xmx[ct] += o.x * o.x + o.y * o.y;
                            }
                        } else {
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const double d = sr * (double)zn[v], xn = (first ? 0.0 : xp[v]) + d;
                                xp[v] = xn;
                                dmx[ct] += d * d;
                                xmx[ct] += xn * xn;
                            }
                        }
                    }
                } else {
                    if (c < nc) *(f4 *)(Z + (size_t)(c0 + c) * ldz + r0 + i) = zn;
                    typedef T16 t4 __attribute__((ext_vector_type(4)));
                    t4 zh, zl;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float zv = PR ? zn[v] * LP_PAIR_Z : zn[v];
                        zh[v] = (T16)zv;
                        if (PR) zl[v] = (T16)(zv - (float)zh[v]);
                    }
                    *(t4 *)(zt + c * zts + i) = zh;
                    if (PR) *(t4 *)(ztl + c * zts + i) = zl;
                }
            }
        if (X && nrm) { // fold the lane sums: lane groups, then waves, then the copy
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
#pragma unroll
                for (int o = 16; o < 64; o <<= 1) {
                    dmx[ct] += __shfl_xor(dmx[ct], o, 64);
                    xmx[ct] += __shfl_xor(xmx[ct], o, 64);
                }
                if (g == 0) {
                    atomicAdd(&xnrm[ct * 16 + m], dmx[ct]);
                    atomicAdd(&xnrm[LV_CW + ct * 16 + m], xmx[ct]);
                }
            }
        }
    }
    if (X && nrm) {
        __syncthreads();
        if (tid < 2 * LV_CW) {
            const int which = tid / LV_CW, c = tid - which * LV_CW;
            if (c < nc)
                atomicAdd(nrm + ((size_t)(blockIdx.x % LP_NRM_COPIES) * ncols + c0 + c) * 2 + which,
                          xnrm[tid]);
        }
    }
    if (a.Rn == 0) return;
    __syncthreads();
    // 4. partial V_{d-1}^T z over the work-group's rows, in the finishing
    // passes' layout (tiles of LV_RB rows of the segment of level d-1)
    {
        const int sn = r0 / a.segn, tile = (r0 - sn * a.segn) / LV_RB, st = sn * a.tps + tile;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int task = wv + 8 * p;
            if (task < nktn * 4) {
                const int ct = task & 3, kt = task >> 2;
                const T16 *pz = zt + (ct * 16 + m) * zts + 8 * g;
                f4 acc = f4{0.f, 0.f, 0.f, 0.f};
                if (PR) {
                    const T16 *pl = ztl + (ct * 16 + m) * zts + 8 * g;
#pragma unroll
                    for (int u = 0; u < NU; ++u)
                        acc = mfma3<T16>(av[p][u], avl[p][u], *(const v8 *)(pz + 32 * u),
                                         *(const v8 *)(pl + 32 * u), acc);
                    acc *= a.isVn / LP_PAIR_Z;
                } else {
#pragma unroll
                    for (int u = 0; u < NU; ++u) acc = Lp<T16>::mfma(av[p][u], *(const v8 *)(pz + 32 * u), acc);
                }
                const int k = kt * 16 + 4 * g, kg = k / LP_KG, kl = k - kg * LP_KG;
                float *pp = part + ((size_t)(st * a.nkg + kg) * a.ncg + blockIdx.y) * LP_BLK +
                            (ct * 16 + m) * LP_KG + kl;
                *(f4 *)pp = acc;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Reassociated upper levels (pair storage, from the FP64 factorization):
//   k_lp_tail   leaves the span's share of V_d^T z for every d < d0,
//   k_lp_up_t0  sums the shares per segment into FP64 T0 of every upper
//               level (rows soff[d] + s*R + k, column stride ws),
//   level steps the FP64 steps of the plain solve (plr_up.hip): W of every
//               upper level from T0 and the cross-level blocks
//               G_{d,e}[s] = V_d|_s^T Ut_e|_s,
//   k_lp_upx    x (+)= max|r| * (z_tail + sum_{d < d0} Ut_d W_d), one pass.
// z is read twice and x once, instead of z read and written per level.

struct LpUpTab
{
    int N, d0, ncols, nspan;
    int R[PLR_MAXL], Rp[PLR_MAXL];
    size_t soff[PLR_MAXL + 1]; // offsets of the FP64 layout (true ranks)
    size_t upoff[PLR_MAXL];
    size_t eoff[PLR_MAXL + 1]; // first output entry of each level (segments x ranks x columns)
};

// T0[(soff[d] + s*R + k) * ws + c]: the sum over the spans of segment s of
// the tail's shares, one thread per entry (columns fastest, so that the
// shares are read and T0 written along the columns), spans summed in a
// fixed order.
__global__ __launch_bounds__(256) void k_lp_up_t0(LpUpTab a, const float *__restrict__ up,
                                                  double *__restrict__ T0, size_t ws)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.eoff[a.d0]) return;
    int d = 0;
    while (e >= a.eoff[d + 1]) ++d;
    const size_t loc = e - a.eoff[d];
    const int n = a.ncols, c = (int)(loc % n);
    const size_t sk = loc / n; // s * R + k
    const int R = a.R[d], Rp = a.Rp[d];
    const int s = (int)(sk / R), k = (int)(sk - (size_t)s * R);
    const int per = a.nspan >> (d + 1); // spans per segment
    const float *p = up + a.upoff[d] + ((size_t)s * per * Rp + k) * n + c;
    const size_t step = (size_t)Rp * n;
    // up to 32 loads in flight per thread (the top level sums 64 spans:
    // two memory waits instead of sixteen)
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    int j = 0;
    for (; j + 32 <= per; j += 32) {
        float v[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) v[u] = p[(size_t)(j + u) * step];
#pragma unroll
        for (int u = 0; u < 32; ++u) acc[u & 3] += v[u];
    }
    for (; j + 4 <= per; j += 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = p[(size_t)(j + u) * step];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += v[u];
    }
    for (; j < per; ++j) acc[j & 3] += p[(size_t)j * step];
    T0[(a.soff[d] + sk) * ws + c] = (double)((acc[0] + acc[1]) + (acc[2] + acc[3]));
}

struct LpUpxTab
{
    int N, d0;
    int R[PLR_MAXL], Rp[PLR_MAXL], seg[PLR_MAXL];
    size_t soff[PLR_MAXL];
    const void *Ut[PLR_MAXL];
    size_t ulo;
    float isU[PLR_MAXL];
};

static constexpr int UX_RB = 256, UX_THREADS = 512, UX_RT = UX_RB / 128, UX_WS = 36;

// LDS of k_lp_upx: W [d0][CW][UX_WS] (floats)
static size_t lpx_lds(int d0) { return (size_t)d0 * LV_CW * UX_WS * sizeof(float); }

// One work-group per (UX_RB rows, 64 right-hand sides), eight waves of
// UX_RT 16-row tiles.  Every level's W for the work-group's segment is
// staged in LDS as FP32 with a power-of-two scale per (level, column), split
// into 16-bit operands (pairs: hi + lo times LP_PAIR_Z) against Ut's rows;
// each level's sums are scaled back before they join z.
template <typename T16, bool PR>
__global__ __launch_bounds__(UX_THREADS) void k_lp_upx(LpUpxTab a, const float *__restrict__ Z,
                                                       int ldz, const double *__restrict__ Wup,
                                                       size_t ws, const double *__restrict__ scale,
                                                       double *__restrict__ X, int ldx, int first,
                                                       int ncols, double *__restrict__ nrm)
{
    typedef typename Lp<T16>::v8 v8;
    extern __shared__ float wl[]; // [d][c][k]
    __shared__ unsigned smax[PLR_MAXL][LV_CW];
    __shared__ double xnrm[2 * LV_CW];
    const int tid = threadIdx.x, lane = tid & 63, m = lane & 15, g = lane >> 4, wv = tid >> 6;
    const int r0 = blockIdx.x * UX_RB, c0 = blockIdx.y * LV_CW;
    const int ri = r0 + wv * 16 * UX_RT;
    const int nc = min(LV_CW, ncols - c0), d0 = a.d0;
    // z_tail as the D operand (rows ri + 16rt + 4g + v of column 16ct + m)
    f4 zo[UX_RT][4];
#pragma unroll
    for (int rt = 0; rt < UX_RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
            zo[rt][ct] = *(const f4 *)(Z + (size_t)(c0 + min(ct * 16 + m, nc - 1)) * ldz + ri + 16 * rt + 4 * g);
    for (int e = tid; e < PLR_MAXL * LV_CW; e += UX_THREADS) (&smax[0][0])[e] = 0u;
    if (tid < 2 * LV_CW) xnrm[tid] = 0.0;
    __syncthreads();
    // W of every level for this segment, ranks past R zero
    for (int d = 0; d < d0; ++d) {
        const int R = a.R[d], s = r0 / a.seg[d];
        const double *wd = Wup + (a.soff[d] + (size_t)s * R) * ws + c0;
        for (int e0 = 0; e0 < 32 * LV_CW; e0 += 4 * UX_THREADS) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * UX_THREADS + tid, k = e / LV_CW, c = e - k * LV_CW;
                v[u] = wd[(size_t)min(k, R - 1) * ws + min(c, nc - 1)];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = e0 + u * UX_THREADS + tid, k = e / LV_CW, c = e - k * LV_CW;
                const float x = (k < R && c < nc) ? (float)v[u] : 0.f;
                wl[(d * LV_CW + c) * UX_WS + k] = x;
                atomicMax(&smax[d][c], __float_as_uint(fabsf(x)));
            }
        }
    }
    __syncthreads();
    f4 tot[UX_RT][4];
#pragma unroll
    for (int rt = 0; rt < UX_RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) tot[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
    // Ut rows of the next level in flight while the current one is used
    auto uload = [&](int d, v8 (*bq)[2]) {
        const T16 *U = (const T16 *)a.Ut[d];
        const int Rp = a.Rp[d];
#pragma unroll
        for (int rt = 0; rt < UX_RT; ++rt) {
            const size_t idx = (size_t)(ri + 16 * rt + m) * Rp + 8 * g;
            bq[rt][0] = *(const v8 *)(U + idx);
            if (PR) bq[rt][1] = *(const v8 *)(U + a.ulo + idx);
        }
    };
    v8 bA[UX_RT][2], bB[UX_RT][2];
    auto level = [&](int d, v8 (*bq)[2]) {
        f4 acc[UX_RT][4];
#pragma unroll
        for (int rt = 0; rt < UX_RT; ++rt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
        float sc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            sc[ct] = lp_pow2(smax[d][ct * 16 + m]);
            const f8 w8 = lds8(wl + (d * LV_CW + ct * 16 + m) * UX_WS + 8 * g);
            if (PR) {
                v8 wh, wlo;
                split8<T16>(w8, LP_PAIR_Z / sc[ct], wh, wlo);
#pragma unroll
                for (int rt = 0; rt < UX_RT; ++rt)
                    acc[rt][ct] = mfma3<T16>(bq[rt][0], bq[rt][1], wh, wlo, acc[rt][ct]);
            } else {
                const v8 bw = cvt8<T16>(w8, 1.f / sc[ct]);
#pragma unroll
                for (int rt = 0; rt < UX_RT; ++rt) acc[rt][ct] = Lp<T16>::mfma(bq[rt][0], bw, acc[rt][ct]);
            }
        }
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const float f = PR ? sc[ct] * (a.isU[d] / LP_PAIR_Z) : sc[ct];
#pragma unroll
            for (int rt = 0; rt < UX_RT; ++rt) tot[rt][ct] += acc[rt][ct] * f;
        }
    };
    uload(0, bA);
    for (int d = 0; d < d0; d += 2) {
        if (d + 1 < d0) uload(d + 1, bB);
        level(d, bA);
        if (d + 1 >= d0) break;
        if (d + 2 < d0) uload(d + 2, bA);
        level(d + 1, bB);
    }
    // x (+)= max|r| * z, with the correction norms
    const bool xal = ((uintptr_t)X & 15) == 0 && ldx % 2 == 0;
    double dmx[4] = {0.0, 0.0, 0.0, 0.0}, xmx[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int rt = 0; rt < UX_RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const int c = ct * 16 + m, i = (wv * UX_RT + rt) * 16 + 4 * g;
            if (c >= nc) continue;
            f4 zn;
#pragma unroll
            for (int v = 0; v < 4; ++v) zn[v] = zo[rt][ct][v] + tot[rt][ct][v];
            double *xp = X + (size_t)(c0 + c) * ldx + r0 + i;
            const double sr = scale[c0 + c] > 0.0 ? scale[c0 + c] : 1.0;
            if (xal) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    double2 o = first ? double2{0.0, 0.0} : *(const double2 *)(xp + 2 * h);
                    const double e0 = sr * (double)zn[2 * h], e1 = sr * (double)zn[2 * h + 1];
                    o.x += e0;
                    o.y += e1;
                    *(double2 *)(xp + 2 * h) = o;
                    dmx[ct] += e0 * e0 + e1 * e1;
                    xmx[ct] += o.x * o.x + o.y * o.y;
                }
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double dv = sr * (double)zn[v], xn = (first ? 0.0 : xp[v]) + dv;
                    xp[v] = xn;
                    dmx[ct] += dv * dv;
                    xmx[ct] += xn * xn;
                }
            }
        }
    if (!nrm) return;
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
#pragma unroll
        for (int o = 16; o < 64; o <<= 1) {
            dmx[ct] += __shfl_xor(dmx[ct], o, 64);
            xmx[ct] += __shfl_xor(xmx[ct], o, 64);
        }
        if (g == 0) {
            atomicAdd(&xnrm[ct * 16 + m], dmx[ct]);
            atomicAdd(&xnrm[LV_CW + ct * 16 + m], xmx[ct]);
        }
    }
    __syncthreads();
    if (tid < 2 * LV_CW) {
        const int which = tid / LV_CW, c = tid - which * LV_CW;
        if (c < nc)
            atomicAdd(nrm + ((size_t)(blockIdx.x % LP_NRM_COPIES) * ncols + c0 + c) * 2 + which,
                      xnrm[tid]);
    }
}

// Host side of the second form.
struct lp_v2
{
    int ok, S, cw, d0, RT, tsegs;
};

static lp_v2 lp_v2_shape(const dplrh_plan *p, const plr_lp *q)
{
    lp_v2 v;
    memset(&v, 0, sizeof(v));
    const int N = p->N, L = p->L, b = p->b;
    if (N == 0 || b % 64 != 0 || b > 512) return v;
    for (int d = 0; d < L; ++d)
        if (q->Rp[d] > 64) return v;
    for (int S = 512; S >= b; S >>= 1) {
        if (S > N) continue;
        int d0 = 0;
        while ((N >> d0) > S) ++d0;
        if ((N >> d0) != S || d0 > L) continue;
        if (S < N && S < 2 * LV_RB) continue; // the upper-level pass: whole row blocks per segment
        int RT = 32;
        for (int d = d0 > 0 ? d0 - 1 : 0; d < L; ++d) RT = q->Rp[d] > RT ? q->Rp[d] : RT;
        const int tsegs = S / b > 2 ? S / b : 2;
        for (int cw = 64; cw >= 32; cw -= 32)
            if (lpt_lds(S, cw, RT, tsegs) + 1024 <= 160 * 1024) { // + static LDS
                v.ok = 1;
                v.S = S;
                v.cw = cw;
                v.d0 = d0;
                v.RT = RT;
                v.tsegs = tsegs;
                return v;
            }
    }
    return v;
}

static bool lp_v2_ok(const dplrh_plan *p, const plr_lp *q) { return lp_v2_shape(p, q).ok != 0; }

// partial blocks the upper levels need (k_lp_level writes, k_lp_finish sums)
static size_t lp_v2_part(const dplrh_plan *p, const plr_lp *q, const lp_v2 &v, int n)
{
    size_t need = 0;
    for (int d = 0; d + 1 < v.d0; ++d) {
        LpTab t;
        lp_tab(p, q, d, d + 1, n, &t, LV_RB);
        if ((size_t)t.part_off[1] > need) need = t.part_off[1];
    }
    return need;
}

template <typename T16, int CW, bool PR>
static int lp_launch_tail(const LpTailTab &tt, size_t lds, dim3 grid, const plr_lp *q, int n,
                          const double *r, int ldr, const double *pm, int pm_n, double *nrm,
                          hipStream_t st)
{
    static bool attr = false;
    if (!attr) {
        // the kernel's static LDS counts against the 160 KB too
        PLR_HIP(hipFuncSetAttribute((const void *)k_lp_tail<T16, CW, PR>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
        attr = true;
    }
    hipLaunchKernelGGL((k_lp_tail<T16, CW, PR>), grid, dim3(LT_THREADS), lds, st, tt, r, ldr, pm, pm_n,
                       q->scale, q->z32, tt.N, q->T, q->t_per_col, n, nrm);
    PLR_HIP(hipGetLastError());
    return 0;
}

// One low-precision solve of the residual r (ldr; pm_n partial maxima of
// |r| per column in pm) added into X:
// X = (first ? 0 : X) + max|r| * solve_lp(r / max|r|).
template <typename T16, bool PR>
static int lp_solve_v2(const dplrh_plan *p, plr_lp *q, const lp_v2 &v, int n, int first,
                       double *X, int ldx, const double *r, int ldr, const double *pm, int pm_n,
                       double *nrm, hipStream_t st)
{
    const int N = p->N, L = p->L;
    const T16 *Vh = (const T16 *)q->Vh, *Uh = (const T16 *)q->Uh;
    dplrh_plan *pm_ = const_cast<dplrh_plan *>(p);
    // the reassociated upper levels: pair storage (taken from the FP64
    // factorization, whose level steps and cross-level blocks they use)
    const bool up = PR && plr_up_ready(p, v.d0);
    const int nspan = N / v.S;
    size_t upoff[PLR_MAXL] = {0}, upn = 0;
    if (up)
        for (int d = 0; d < v.d0; ++d) {
            upoff[d] = upn;
            upn += (size_t)n * nspan * q->Rp[d];
        }
    int rc = lp_part(p, q, up ? (upn + LP_BLK - 1) / LP_BLK : lp_v2_part(p, q, v, n));
    if (rc) return rc;
    if (up && (rc = plr_up_reserve(pm_, v.d0, n))) return rc;
    LpTailTab tt;
    memset(&tt, 0, sizeof(tt));
    tt.N = N;
    tt.b = p->b;
    tt.L = L;
    tt.d0 = v.d0;
    tt.span = v.S;
    tt.RT = v.RT;
    tt.tsegs = v.tsegs;
    for (int d = 0; d < L; ++d) {
        tt.Rp[d] = q->Rp[d];
        tt.V[d] = Vh + q->off[d];
        tt.Ut[d] = Uh + q->off[d];
        tt.S[d] = q->Sf + q->soff[d];
        tt.K32[d] = d >= 2 ? q->K32 + q->koff[d] : nullptr;
        tt.K64[d] = d < 2 ? q->K64 + q->koff[d] : nullptr;
    }
    tt.D = q->Dh;
    tt.vlo = tt.ulo = q->off[L];
    tt.dlo = (size_t)N * p->b;
    for (int d = 0; d < L; ++d) {
        tt.isV[d] = q->isV[d];
        tt.isU[d] = q->isU[d];
    }
    tt.isD = q->isD;
    if (up) {
        tt.upart = q->part;
        tt.nspan = nspan;
        memcpy(tt.upoff, upoff, sizeof(upoff));
    }
    const size_t lds = lpt_lds(v.S, v.cw, v.RT, v.tsegs);
    const dim3 grid(N / v.S, (n + v.cw - 1) / v.cw);
    rc = v.cw == 64 ? lp_launch_tail<T16, 64, PR>(tt, lds, grid, q, n, r, ldr, pm, pm_n, nrm, st)
                    : lp_launch_tail<T16, 32, PR>(tt, lds, grid, q, n, r, ldr, pm, pm_n, nrm, st);
    if (rc) return rc;
    if (v.d0 == 0) {
        hipLaunchKernelGGL(k_lp_update, dim3((N + 255) / 256, n), dim3(256), 0, st, q->z32, N,
                           q->scale, X, ldx, n, first, nrm);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    if (up) {
        LpUpTab ut;
        memset(&ut, 0, sizeof(ut));
        ut.N = N;
        ut.d0 = v.d0;
        ut.ncols = n;
        ut.nspan = nspan;
        size_t e = 0;
        for (int d = 0; d < v.d0; ++d) {
            ut.R[d] = p->R[d];
            ut.Rp[d] = q->Rp[d];
            ut.soff[d] = p->soff[d];
            ut.upoff[d] = upoff[d];
            ut.eoff[d] = e;
            e += ((size_t)2 << d) * p->R[d] * n;
        }
        ut.eoff[v.d0] = e;
        hipLaunchKernelGGL(k_lp_up_t0, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, st, ut,
                           q->part, p->sw.uT0, (size_t)p->sw.ucols);
        PLR_HIP(hipGetLastError());
        if ((rc = plr_up_steps(pm_, v.d0, n, st))) return rc < 0 ? rc : -10;
        LpUpxTab xt;
        memset(&xt, 0, sizeof(xt));
        xt.N = N;
        xt.d0 = v.d0;
        for (int d = 0; d < v.d0; ++d) {
            xt.R[d] = p->R[d];
            xt.Rp[d] = q->Rp[d];
            xt.seg[d] = N >> (d + 1);
            xt.soff[d] = p->soff[d];
            xt.Ut[d] = Uh + q->off[d];
            xt.isU[d] = q->isU[d];
        }
        xt.ulo = q->off[L];
        static bool xattr = false;
        if (!xattr) {
            PLR_HIP(hipFuncSetAttribute((const void *)k_lp_upx<T16, PR>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
            xattr = true;
        }
        hipLaunchKernelGGL((k_lp_upx<T16, PR>), dim3(N / UX_RB, (n + LV_CW - 1) / LV_CW),
                           dim3(UX_THREADS), lpx_lds(v.d0), st, xt, q->z32, N, p->sw.uW,
                           (size_t)p->sw.ucols, q->scale, X, ldx, first, n, nrm);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    static bool attr = false;
    if (!attr) {
        // lpl_lds(64) is 92 KB; the kernel's static LDS (1.25 KB) counts
        // against the 160 KB too
        PLR_HIP(hipFuncSetAttribute((const void *)k_lp_level<T16, 32, PR>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024));
        PLR_HIP(hipFuncSetAttribute((const void *)k_lp_level<T16, 64, PR>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024));
        attr = true;
    }
    const int ncg = (n + LV_CW - 1) / LV_CW;
    for (int d = v.d0 - 1; d >= 0; --d) {
        LpLvlTab lt;
        memset(&lt, 0, sizeof(lt));
        lt.N = N;
        lt.R = q->Rp[d];
        lt.Rn = d > 0 ? q->Rp[d - 1] : 0;
        lt.seg = N >> (d + 1);
        lt.segn = N >> d;
        lt.tps = lt.segn / LV_RB;
        lt.nkg = lt.Rn / LP_KG;
        lt.ncg = ncg;
        lt.Ut = Uh + q->off[d];
        lt.Vn = d > 0 ? Vh + q->off[d - 1] : nullptr;
        lt.S = q->Sf + q->soff[d];
        lt.K32 = d >= 2 ? q->K32 + q->koff[d] : nullptr;
        lt.K64 = d < 2 ? q->K64 + q->koff[d] : nullptr;
        lt.vlo = lt.ulo = q->off[L];
        lt.isU = q->isU[d];
        lt.isVn = d > 0 ? q->isV[d - 1] : 1.f;
        if (lt.R <= 32 && lt.Rn <= 32)
            hipLaunchKernelGGL((k_lp_level<T16, 32, PR>), dim3(N / LV_RB, ncg), dim3(LV_THREADS),
                               lpl_lds(lt.R, PR), st, lt, q->T, q->t_per_col, q->z32, N, q->part,
                               q->scale, d == 0 ? X : nullptr, ldx, first, n, nrm);
        else
            hipLaunchKernelGGL((k_lp_level<T16, 64, PR>), dim3(N / LV_RB, ncg), dim3(LV_THREADS),
                               lpl_lds(lt.R, PR), st, lt, q->T, q->t_per_col, q->z32, N, q->part,
                               q->scale, d == 0 ? X : nullptr, ldx, first, n, nrm);
        PLR_HIP(hipGetLastError());
        if (d > 0 && lt.Rn > 0) {
            LpTab t;
            lp_tab(p, q, d - 1, d, n, &t, LV_RB);
            hipLaunchKernelGGL(k_lp_finish1, dim3(t.fin_off[1] * (LP_BLK / (256 / lpf_phases(t.tps[0])))),
                               dim3(256), 0, st, t, q->part, q->T, q->t_per_col);
            PLR_HIP(hipGetLastError());
        }
    }
    return 0;
}

static int lp_solve_v2_any(const dplrh_plan *p, plr_lp *q, int n, int first, double *X, int ldx,
                           const double *r, int ldr, const double *pm, int pm_n, double *nrm,
                           hipStream_t st)
{
    const lp_v2 v = lp_v2_shape(p, q);
    if (!v.ok) return 1; // shape not covered: the caller takes the first form
    switch (q->type) {
    case 1: return lp_solve_v2<_Float16, false>(p, q, v, n, first, X, ldx, r, ldr, pm, pm_n, nrm, st);
    case 2: return lp_solve_v2<__bf16, false>(p, q, v, n, first, X, ldx, r, ldr, pm, pm_n, nrm, st);
    default: return lp_solve_v2<_Float16, true>(p, q, v, n, first, X, ldx, r, ldr, pm, pm_n, nrm, st);
    }
}
