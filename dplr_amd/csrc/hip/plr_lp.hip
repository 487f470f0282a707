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
    __shared__ double red[256];
    const int c = blockIdx.y, ch = blockIdx.x;
    const int len = (N + LP_SCALE_CHUNKS - 1) / LP_SCALE_CHUNKS;
    const int beg = ch * len, end = min(N, beg + len);
    const double *rc = r + (size_t)c * ldr;
    double mv = 0.0;
    for (int i = beg + threadIdx.x; i < end; i += 256) mv = fmax(mv, fabs(rc[i]));
    red[threadIdx.x] = mv;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) pmax[(size_t)c * LP_SCALE_CHUNKS + ch] = red[0];
}

__global__ __launch_bounds__(256) void k_lp_scale_rhs(const double *r, int ldr, int N,
                                                      const double *pmax, double *scale, float *r32)
{
    const int c = blockIdx.y;
    double mv = 0.0;
    for (int j = 0; j < LP_SCALE_CHUNKS; ++j) mv = fmax(mv, pmax[(size_t)c * LP_SCALE_CHUNKS + j]);
    if (blockIdx.x == 0 && threadIdx.x == 0) scale[c] = mv;
    const double inv = 1.0 / (mv > 0.0 ? mv : 1.0);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N) r32[(size_t)c * N + i] = (float)(r[(size_t)c * ldr + i] * inv);
}

// x[c] (+)= scale[c] * z[c]  (FP64 accumulation of the correction; the
// first correction sets x)
__global__ void k_lp_update(const float *z, int N, const double *scale, double *x, int ldx, int n,
                            int first)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * n) return;
    const int c = (int)(e / N), i = (int)(e % N);
    const double sc = scale[c] > 0.0 ? scale[c] : 1.0;
    double *xe = x + (size_t)c * ldx + i;
    *xe = (first ? 0.0 : *xe) + sc * (double)z[e];
}

// ---------------------------------------------------------------------------
// host side

struct plr_lp
{
    int type; // 1 = f16, 2 = bf16
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
};

static void lp_free(plr_lp *q)
{
    if (!q) return;
    void *b[] = {q->Vh, q->Uh, q->Dh, q->Sf, q->K64, q->K32, q->T, q->W, q->wscale, q->part,
                 q->r32, q->z32, q->r, q->scale, q->pmax};
    for (void *x : b)
        if (x) (void)hipFree(x);
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

static void lp_tab(const dplrh_plan *p, const plr_lp *q, int d0, int d1, int ncols, LpTab *t)
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
        const int T = seg < LP_TMAX ? seg : LP_TMAX, tps = seg / T, nkg = R / LP_KG;
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
    void *old[] = {q->T, q->W, q->wscale, q->r32, q->z32, q->r, q->scale, q->pmax};
    for (void *x : old)
        if (x) (void)hipFree(x);
    q->T = q->W = q->wscale = q->r32 = q->z32 = nullptr;
    q->r = q->scale = q->pmax = nullptr;
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

// type: 1 = f16, 2 = bf16.  Requires plr_mm_ok(p).
extern "C" int dplrh_lp_factor(dplrh_plan *p, int type, int *info, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    *info = 0;
    if (!plr_mm_ok(p)) {
        plr_set_error("mixed precision needs the leaf order to be a multiple of 64");
        return -5;
    }
    if (type != 1 && type != 2) {
        plr_set_error("unknown low-precision type %d", type);
        return -5;
    }
    plr_lp_release(p);
    ++p->ws_gen;
    plr_lp *q = new plr_lp;
    memset(q, 0, sizeof(*q));
    q->type = type;
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
    if ((rc = alloc(&q->Vh, off * 2)) || (rc = alloc(&q->Uh, off * 2)) ||
        (rc = alloc(&q->Dh, (size_t)p->N * p->b * 2)) || (rc = alloc((void **)&q->Sf, so * 4)) ||
        (rc = alloc((void **)&q->K64, k64 * 8)) || (rc = alloc((void **)&q->K32, k32 * 4))) {
        lp_free(q);
        return rc;
    }
    g_lp.push_back({p, q});
    rc = type == 1 ? lp_factor_t<_Float16>(p, q, info, st) : lp_factor_t<__bf16>(p, q, info, st);
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

// X := A^{-1} B by mixed-precision solves refined with FP64 residuals:
// `iters` solves (iters-1 residuals).  When res != NULL, res[c] receives
// max|b - A x| of column c after the last solve (one more FP64 residual).
static int lp_solve_body(dplrh_plan *p, plr_lp *q, int n, const double *dB, int ldb, double *dX,
                         int ldx, int iters, double *res, hipStream_t st)
{
    const int N = p->N;
    if (N == 0 || n == 0) return 0;
    int rc = 0;
    const int chunk = n < 256 ? n : 256;
    if ((rc = lp_ws(p, q, chunk))) return rc;
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int nc = n - c0 < chunk ? n - c0 : chunk;
        const double *B = dB + (size_t)c0 * ldb;
        double *X = dX + (size_t)c0 * ldx;
        for (int it = 0; it < iters + (res ? 1 : 0); ++it) {
            // r = b - A x (x = 0 on the first pass)
            PLR_HIP(hipMemcpy2DAsync(q->r, (size_t)N * sizeof(double), B, (size_t)ldb * sizeof(double),
                                     (size_t)N * sizeof(double), nc, hipMemcpyDeviceToDevice, st));
            if (it > 0 && (rc = plr_gemm_direct(p, 0, nc, -1.0, 1.0, X, ldx, q->r, N, st))) return rc;
            hipLaunchKernelGGL(k_lp_colmax, dim3(LP_SCALE_CHUNKS, nc), dim3(256), 0, st, q->r, N, N,
                               q->pmax);
            hipLaunchKernelGGL(k_lp_scale_rhs, dim3((N + 255) / 256, nc), dim3(256), 0, st, q->r, N,
                               N, q->pmax, q->scale, q->r32);
            PLR_HIP(hipGetLastError());
            if (it == iters) break; // residual of the final iterate only
            rc = q->type == 1 ? lp_solve_t<_Float16>(p, q, nc, st) : lp_solve_t<__bf16>(p, q, nc, st);
            if (rc) return rc;
            const size_t ne = (size_t)N * nc;
            hipLaunchKernelGGL(k_lp_update, dim3((ne + 255) / 256), dim3(256), 0, st, q->z32, N,
                               q->scale, X, ldx, nc, it == 0);
            PLR_HIP(hipGetLastError());
        }
        if (res) {
            PLR_HIP(hipMemcpyAsync(res + c0, q->scale, nc * sizeof(double), hipMemcpyDeviceToHost, st));
            PLR_HIP(hipStreamSynchronize(st));
        }
    }
    return 0;
}

extern "C" int dplrh_lp_solve(dplrh_plan *p, int n, const double *dB, int ldb, double *dX,
                              int ldx, int iters, double *res, void *stream)
{
    plr_lp *q = lp_of(p);
    if (!q) {
        plr_set_error("solve before a successful mixed-precision factorization");
        return -4;
    }
    if (res) // reads results back to the host: not capturable
        return lp_solve_body(p, q, n, dB, ldb, dX, ldx, iters, res, (hipStream_t)stream);
    const unsigned long long key[10] = {4, (unsigned long long)n, (unsigned long long)dB,
                                        (unsigned long long)ldb, (unsigned long long)dX,
                                        (unsigned long long)ldx, (unsigned long long)iters,
                                        (unsigned long long)q};
    return plr_run_cached(p, key, (hipStream_t)stream, [&](hipStream_t s) {
        return lp_solve_body(p, q, n, dB, ldb, dX, ldx, iters, nullptr, s);
    });
}
