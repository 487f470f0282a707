// Factored FP64 inverse of a PLR matrix and its application to right-hand
// sides (the solve that replaces "dplrgeinv, then apply the explicit
// inverse", reference src/dplrinv.c:93 with src/dplrpmv.c / dplrpdmm.c).
//
// The reference forms B = inv(A) explicitly, node by node in post-order,
// using X = [inv(S1) V1^T u2; V2^T u1 inv(S2)] (src/dplrinv.c:283-315).  The
// same Woodbury identity is used here in product form, level by level:
//
//   inv(A) = F_0 F_1 ... F_{L-1} inv(D_leaf),   F_d = I - Ut_d S_d inv(K_d) Z_d^T
//
// where Ut_d is U_d with every level below d already inverted, and per node
// K = I + Z^T inv(D) W S = [I, M_bot S_bot; M_top S_top, I] with
// M_s = V_d|_s^T Ut_d|_s.  K needs no 1/sig, so sig = 0 or tiny sig cannot
// overflow it (the reference's X holds 1/sig).  Factorization, bottom-up:
//   1. LU with partial pivoting of every leaf, then its inverse (Dinv);
//   2. Ut = Dinv applied to the U columns of all levels;
//   3. for d = L-1 .. 0: M = V_d^T [Ut_{<d} | Ut_d] per segment, build and
//      invert K per node, then Ut_{<d} -= Ut_d S_d inv(K) Z_d^T Ut_{<d}.
// Solve: z = Dinv b, then for d = L-1 .. 0: z -= Ut_d S_d inv(K) Z_d^T z.
// Transpose solve: the factors in reverse order, transposed.
#include <cstdio>
#include <cstring>
#include <vector>

#include "plr_cap.h"
#include "plr_common.h"

// ---------------------------------------------------------------------------
// batched small dense kernels

// In-place LU with partial pivoting of n x n column-major matrices
// A + k*stride, one work-group per matrix.  piv: 0-based pivot rows.
// info[k] = first j (1-based) with an exactly zero pivot, else 0.
__global__ __launch_bounds__(256) void k_lu_batched(double *A, int n, size_t stride, int *piv,
                                                    int *info)
{
    __shared__ double sval[256];
    __shared__ int sidx[256];
    __shared__ int s_p, s_info;
    const int tid = threadIdx.x;
    double *a = A + (size_t)blockIdx.x * stride;
    int *pv = piv + (size_t)blockIdx.x * n;
    if (tid == 0) s_info = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        double best = -1.0;
        int bi = n;
        for (int i = j + tid; i < n; i += blockDim.x) {
            const double v = fabs(a[(size_t)j * n + i]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        sval[tid] = best;
        sidx[tid] = bi;
        __syncthreads();
        for (int off = 128; off > 0; off >>= 1) {
            if (tid < off) {
                const double v2 = sval[tid + off];
                const int i2 = sidx[tid + off];
                if (v2 > sval[tid] || (v2 == sval[tid] && i2 < sidx[tid])) {
                    sval[tid] = v2;
                    sidx[tid] = i2;
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            int p = sidx[0];
            if (p >= n) p = j; // column of NaNs: keep the diagonal
            s_p = p;
            pv[j] = p;
            if (a[(size_t)j * n + p] == 0.0 && s_info == 0) s_info = j + 1;
        }
        __syncthreads();
        const int p = s_p;
        const double pval = a[(size_t)j * n + p];
        if (pval != 0.0) {
            if (p != j)
                for (int c = tid; c < n; c += blockDim.x) {
                    const double t = a[(size_t)c * n + j];
                    a[(size_t)c * n + j] = a[(size_t)c * n + p];
                    a[(size_t)c * n + p] = t;
                }
            __syncthreads();
            const double inv = 1.0 / a[(size_t)j * n + j];
            for (int i = j + 1 + tid; i < n; i += blockDim.x) a[(size_t)j * n + i] *= inv;
            __syncthreads();
            const int m = n - j - 1;
            for (int e = tid; e < m * m; e += blockDim.x) {
                const int ii = j + 1 + e % m, cc = j + 1 + e / m;
                a[(size_t)cc * n + ii] -= a[(size_t)j * n + ii] * a[(size_t)cc * n + j];
            }
        }
        __syncthreads();
    }
    if (tid == 0) info[blockIdx.x] = s_info;
}

// inv(A) from its LU factors, one work-group per matrix, one thread per
// column of the identity: P e_c, then unit-lower and upper substitution.
__global__ __launch_bounds__(256) void k_lu_inverse(const double *LU, const int *piv, int n,
                                                    size_t stride, double *Inv, size_t istride)
{
    const double *a = LU + (size_t)blockIdx.x * stride;
    const int *pv = piv + (size_t)blockIdx.x * n;
    double *out = Inv + (size_t)blockIdx.x * istride;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        double *x = out + (size_t)c * n;
        for (int i = 0; i < n; ++i) x[i] = (i == c) ? 1.0 : 0.0;
        for (int j = 0; j < n; ++j) {
            const int p = pv[j];
            if (p != j) {
                const double t = x[j];
                x[j] = x[p];
                x[p] = t;
            }
        }
        for (int i = 0; i < n; ++i) {
            double s = x[i];
            for (int k = 0; k < i; ++k) s -= a[(size_t)k * n + i] * x[k];
            x[i] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = x[i];
            for (int k = i + 1; k < n; ++k) s -= a[(size_t)k * n + i] * x[k];
            x[i] = s / a[(size_t)i * n + i];
        }
    }
}

// In-place Gauss-Jordan inverse with partial pivoting in LDS, one
// work-group per n x n column-major matrix (A + k*stride -> Inv + k*istride).
// Step k picks the first row i >= k of largest |A[i][k]|: the same pivot
// sequence as LU with partial pivoting, so info[k] is dgetrf's status (first
// exactly-zero pivot, 1-based, or 0).  Row interchanges are undone as column
// interchanges in reverse order at the end.
__global__ __launch_bounds__(256) void k_gj_inverse(const double *A, int n, size_t stride,
                                                    double *Inv, size_t istride, int *info)
{
    extern __shared__ double a[]; // n*n, column-major
    __shared__ double sval[256];
    __shared__ int sidx[256];
    __shared__ int spiv[256];
    __shared__ int s_info;
    const int tid = threadIdx.x, nt = blockDim.x;
    const double *src = A + (size_t)blockIdx.x * stride;
    for (int e = tid; e < n * n; e += nt) a[e] = src[e];
    if (tid == 0) s_info = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += nt) {
            const double v = fabs(a[(size_t)k * n + i]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        sval[tid] = best;
        sidx[tid] = bi;
        __syncthreads();
        for (int off = nt >> 1; off > 0; off >>= 1) {
            if (tid < off) {
                const double v2 = sval[tid + off];
                const int i2 = sidx[tid + off];
                if (v2 > sval[tid] || (v2 == sval[tid] && i2 < sidx[tid])) {
                    sval[tid] = v2;
                    sidx[tid] = i2;
                }
            }
            __syncthreads();
        }
        int p = sidx[0];
        if (p >= n) p = k;
        const double pv = a[(size_t)k * n + p];
        if (tid == 0) {
            spiv[k] = p;
            if (pv == 0.0 && s_info == 0) s_info = k + 1;
        }
        if (p != k)
            for (int j = tid; j < n; j += nt) {
                const double t = a[(size_t)j * n + k];
                a[(size_t)j * n + k] = a[(size_t)j * n + p];
                a[(size_t)j * n + p] = t;
            }
        __syncthreads();
        const double inv = 1.0 / a[(size_t)k * n + k];
        // row k scaled (its diagonal becomes inv below)
        for (int j = tid; j < n; j += nt)
            if (j != k) a[(size_t)j * n + k] *= inv;
        __syncthreads();
        // eliminate column k from every other row; element e = j*n + i,
        // its row and column stepped along (no division per element)
        {
            const int di = nt % n, dj = nt / n;
            int i = tid % n, j = tid / n;
            for (int e = tid; e < n * n; e += nt) {
                if (i != k && j != k) a[(size_t)j * n + i] -= a[(size_t)k * n + i] * a[(size_t)j * n + k];
                i += di;
                j += dj;
                if (i >= n) {
                    i -= n;
                    ++j;
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < n; i += nt)
            a[(size_t)k * n + i] = (i == k) ? inv : -a[(size_t)k * n + i] * inv;
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        const int p = spiv[k];
        if (p != k)
            for (int i = tid; i < n; i += nt) {
                const double t = a[(size_t)k * n + i];
                a[(size_t)k * n + i] = a[(size_t)p * n + i];
                a[(size_t)p * n + i] = t;
            }
        __syncthreads();
    }
    double *dst = Inv + (size_t)blockIdx.x * istride;
    for (int e = tid; e < n * n; e += nt) dst[e] = a[e];
    if (tid == 0) info[blockIdx.x] = s_info;
}

// Register Gauss-Jordan for orders n <= NM: one wavefront per matrix, lane
// i holding row i of the matrix in NM registers (rows and columns past n
// are those of the identity, which the steps leave alone).  Rows are not
// moved: each lane carries its logical position, and step k takes as pivot
// the lane of largest |A[.][k]| among positions >= k (ties: the smaller
// position), exchanging positions with the lane at k -- the pivot sequence
// of LU with partial pivoting, so info is dgetrf's.  The pivot row reaches
// the other lanes by lane reads.  Columns rotate through the registers,
// eight at a time, so the step's column is always in one of eight fixed
// registers.  At the end lane L holds row pos[L] of inv(PA), and column k
// of inv(PA) is column perm[k] of inv(A), perm[k] the lane at position k.
__device__ __forceinline__ double gj_rdlane(double x, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}

// Pivot of step k: the position (uniform) of the largest |a| over the lanes
// whose position is >= k, ties to the smaller position, NaN below every
// number.  The order is that of one 64-bit key per lane (0: position < k,
// 1: NaN, else the bits of |a| plus 2, which order as the values do); the
// lanes of each 16-lane row meet through DPP moves (quad exchanges, then
// rotations by 4 and 8), the four row winners through lane reads.
template <int CTRL>
__device__ __forceinline__ void gj_amax_dpp(unsigned &hi, unsigned &lo, int &p)
{
    const unsigned h2 = (unsigned)__builtin_amdgcn_update_dpp(0, (int)hi, CTRL, 0xF, 0xF, false);
    const unsigned l2 = (unsigned)__builtin_amdgcn_update_dpp(0, (int)lo, CTRL, 0xF, 0xF, false);
    const int p2 = __builtin_amdgcn_update_dpp(0, p, CTRL, 0xF, 0xF, false);
    const bool take = h2 > hi || (h2 == hi && (l2 > lo || (l2 == lo && p2 < p)));
    hi = take ? h2 : hi;
    lo = take ? l2 : lo;
    p = take ? p2 : p;
}

// ok: the row may pivot step k -- a row of the matrix for k < n, a row of
// the identity padding for k >= n (else an all-NaN column would take a
// padding row, which would then leave the n x n block)
__device__ __forceinline__ unsigned long long gj_key(double a, int pos, int k, bool ok)
{
    if (pos < k || !ok) return 0ull;
    return a != a ? 1ull : (unsigned long long)__double_as_longlong(fabs(a)) + 2ull;
}

// the largest of the wave's 32-bit values: the lanes of each 16-lane row
// meet through DPP moves, the four row maxima through lane reads
__device__ __forceinline__ unsigned gj_wave_max(unsigned v)
{
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));
    return max(max((unsigned)__builtin_amdgcn_readlane(v, 0), (unsigned)__builtin_amdgcn_readlane(v, 16)),
               max((unsigned)__builtin_amdgcn_readlane(v, 32), (unsigned)__builtin_amdgcn_readlane(v, 48)));
}

// the position of the wave's largest key (ties: the smaller position).
// When one lane alone holds the largest upper 32 bits, its key is the
// largest; otherwise (equal exponents and leading 20 mantissa bits, or no
// number above zero) the full 64-bit keys and positions are compared.
__device__ __forceinline__ int gj_pivot_key(unsigned long long key, int pos)
{
    unsigned hi = (unsigned)(key >> 32), lo = (unsigned)key;
    const unsigned long long top = __ballot(hi == gj_wave_max(hi));
    if (__popcll(top) == 1) return __builtin_amdgcn_readlane(pos, __ffsll(top) - 1);
    int p = pos;
    gj_amax_dpp<0xB1>(hi, lo, p);  // quad_perm 1,0,3,2
    gj_amax_dpp<0x4E>(hi, lo, p);  // quad_perm 2,3,0,1
    gj_amax_dpp<0x124>(hi, lo, p); // row_ror 4
    gj_amax_dpp<0x128>(hi, lo, p); // row_ror 8
    unsigned bh = __builtin_amdgcn_readlane(hi, 0), bl = __builtin_amdgcn_readlane(lo, 0);
    int bp = __builtin_amdgcn_readlane(p, 0);
#pragma unroll
    for (int r = 1; r < 4; ++r) {
        const unsigned h = __builtin_amdgcn_readlane(hi, 16 * r);
        const unsigned l = __builtin_amdgcn_readlane(lo, 16 * r);
        const int q = __builtin_amdgcn_readlane(p, 16 * r);
        if (h > bh || (h == bh && (l > bl || (l == bl && q < bp)))) {
            bh = h;
            bl = l;
            bp = q;
        }
    }
    return bp;
}

__device__ __forceinline__ int gj_pivot(double a, int pos, int k, bool ok)
{
    return gj_pivot_key(gj_key(a, pos, k, ok), pos);
}

// the same over two rows per lane: the lane's better candidate first
__device__ __forceinline__ int gj_pivot2(double a0, int pos0, bool ok0, double a1, int pos1,
                                         bool ok1, int k)
{
    const unsigned long long k0 = gj_key(a0, pos0, k, ok0), k1 = gj_key(a1, pos1, k, ok1);
    const bool second = k1 > k0 || (k1 == k0 && pos1 < pos0);
    return gj_pivot_key(second ? k1 : k0, second ? pos1 : pos0);
}

template <int NM>
__global__ __launch_bounds__(256) void k_gj_inverse_w(const double *A, int n, size_t stride,
                                                      double *Inv, size_t istride, int *info,
                                                      int count)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int mat = blockIdx.x * 4 + wv;
    if (mat >= count) return; // the whole wavefront leaves together
    const double *src = A + (size_t)mat * stride;
    double a[NM];
#pragma unroll
    for (int j = 0; j < NM; ++j)
        a[j] = (lane < n && j < n) ? src[(size_t)j * n + lane] : (j == lane ? 1.0 : 0.0);
    int pos = lane, sinfo = 0;
    const int steps = (n + 7) & ~7;
    for (int k0 = 0; k0 < steps; k0 += 8) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int k = k0 + c;
            // pivot: largest |a[c]| over positions >= k, ties to the smaller
            // position (NaN ranks below every number: one order for all lanes)
            const int key = gj_pivot(a[c], pos, k, (lane < n) == (k < n));
            const bool piv = pos == key;
            const int pl = __ffsll((unsigned long long)__ballot(piv)) - 1;
            if (pos == k)
                pos = key;
            else if (piv)
                pos = k;
            const double pv = gj_rdlane(a[c], pl);
            if (pv == 0.0 && sinfo == 0) sinfo = k + 1;
            const double inv = 1.0 / pv;
            const double f = a[c];
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                if (j == c) continue;
                const double pj = gj_rdlane(a[j], pl) * inv;
                a[j] = piv ? pj : a[j] - f * pj;
            }
            a[c] = piv ? inv : -f * inv;
        }
        // rotate eight columns: the next block's columns into a[0..7]
        double t[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = a[j];
#pragma unroll
        for (int j = 0; j < NM - 8; ++j) a[j] = a[j + 8];
#pragma unroll
        for (int j = 0; j < 8; ++j) a[NM - 8 + j] = t[j];
    }
    // register j now holds column (j + steps) % NM; column col of inv(PA)
    // is column perm[col] of inv(A), perm[col] the lane at position col
    double *dst = Inv + (size_t)mat * istride;
#pragma unroll
    for (int j = 0; j < NM; ++j) {
        const int col = (j + steps) % NM;
        const int pc = __ffsll((unsigned long long)__ballot(pos == col)) - 1;
        if (col < n && lane < n) dst[(size_t)pc * n + pos] = a[j];
    }
    if (lane == 0) info[mat] = sinfo;
}

// The same elimination for few matrices: one work-group of NM/CW waves per
// matrix, wave w holding columns CW*w .. CW*w + CW - 1 (lane i: row i).  The pivot
// of step k + 1 is found one step early: after step k's pivot is known,
// the wave that owns column k + 1 updates that column first, finds its
// pivot (same rule and positions as k_gj_inverse_w) and publishes the
// column, the pivot's position, lane and value through LDS, double-buffered,
// before it updates its other columns; so one barrier per step suffices
// and the owner's other updates are off the step's critical path.  Output
// and info as k_gj_inverse_w.
// columns per wave of k_gj_inverse_q: at order 64 eight waves of 8 columns
// beat four of 16 and sixteen of 4 (less update work for the wave that
// finds the pivot, against one more barrier participant per halving)
static constexpr int GJQ_CW = 8;
template <int NM, int CW>
__global__ __launch_bounds__(NM / CW * 64) void k_gj_inverse_q(const double *A, int n, size_t stride,
                                                         double *Inv, size_t istride, int *info)
{
    __shared__ double colv[2][64];
    __shared__ double pval[2]; // the pivot
    __shared__ int pk[2][2];   // position of the pivot, its lane
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mat = blockIdx.x;
    const double *src = A + (size_t)mat * stride;
    double a[CW];
#pragma unroll
    for (int j = 0; j < CW; ++j) {
        const int gc = CW * wv + j;
        a[j] = (lane < n && gc < n) ? src[(size_t)gc * n + lane] : (gc == lane ? 1.0 : 0.0);
    }
    int pos = lane, sinfo = 0;
    const int nb = min(NM, (n + CW - 1) / CW * CW) / CW, steps = CW * nb;
    // publish the pivot of step k on column value x (positions >= k, ties
    // to the smaller position, NaN ranks below every number)
    auto publish = [&](int k, double x) {
        const int key = gj_pivot(x, pos, k, (lane < n) == (k < n));
        const int pl = __ffsll((unsigned long long)__ballot(pos == key)) - 1;
        const double pv = gj_rdlane(x, pl);
        colv[k & 1][lane] = x;
        if (lane == 0) {
            pk[k & 1][0] = key;
            pk[k & 1][1] = pl;
            pval[k & 1] = pv;
        }
    };
    if (wv == 0) publish(0, a[0]);
    for (int kb = 0; kb < nb; ++kb) {
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int k = CW * kb + c, buf = k & 1;
            __syncthreads();
            const int key = pk[buf][0], pl = pk[buf][1];
            const bool piv = pos == key;
            if (pos == k)
                pos = key;
            else if (piv)
                pos = k;
            const double pv = pval[buf];
            if (pv == 0.0 && sinfo == 0) sinfo = k + 1;
            const double inv = 1.0 / pv, f = colv[buf][lane];
            // column k + 1: register c1 of wave kb (c < CW - 1) or kb + 1
            const int c1 = (c + 1) % CW;
            const bool next = k + 1 < steps && wv == kb + (c == CW - 1);
            if (next) {
                const double pj = gj_rdlane(a[c1], pl) * inv;
                a[c1] = piv ? pj : a[c1] - f * pj;
                publish(k + 1, a[c1]);
            }
#pragma unroll
            for (int j = 0; j < CW; ++j) {
                if ((wv == kb && j == c) || (next && j == c1)) continue;
                const double pj = gj_rdlane(a[j], pl) * inv;
                a[j] = piv ? pj : a[j] - f * pj;
            }
            if (wv == kb) a[c] = piv ? inv : -f * inv;
        }
    }
    // column col of inv(PA) is column perm[col] of inv(A), perm[col] the
    // lane at position col
    double *dst = Inv + (size_t)mat * istride;
#pragma unroll
    for (int j = 0; j < CW; ++j) {
        const int col = CW * wv + j;
        const int pc = __ffsll((unsigned long long)__ballot(pos == col)) - 1;
        if (col < n && lane < n) dst[(size_t)pc * n + pos] = a[j];
    }
    if (threadIdx.x == 0) info[mat] = sinfo;
}

// The elimination of k_gj_inverse_q for orders up to 128: one work-group
// of eight waves per matrix, wave w holding columns 16w .. 16w+15 of rows
// l and 64 + l in lane l (two registers per column).  The owner of the
// step's column finds the pivot over both rows of every lane and publishes
// the column and the pivot's position through LDS (double-buffered, one
// barrier per step); the pivot row's entries come from lane reads of the
// half that holds it.  Pivot rule, arithmetic per element, output and info
// as k_gj_inverse_w.
__global__ __launch_bounds__(512) void k_gj_inverse_r(const double *A, int n, size_t stride,
                                                      double *Inv, size_t istride, int *info)
{
    __shared__ double colv[2][128];
    __shared__ int pk[2];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mat = blockIdx.x;
    const double *src = A + (size_t)mat * stride;
    double a[16][2];
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int gc = 16 * wv + j, row = 64 * h + lane;
            a[j][h] = (row < n && gc < n) ? src[(size_t)gc * n + row] : (gc == row ? 1.0 : 0.0);
        }
    int pos0 = lane, pos1 = 64 + lane, sinfo = 0;
    const int nb = min(128, (n + 15) & ~15) / 16;
    for (int kb = 0; kb < nb; ++kb) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int k = 16 * kb + c, buf = k & 1;
            if (wv == kb) {
                const int key = gj_pivot2(a[c][0], pos0, (lane < n) == (k < n), a[c][1], pos1,
                                          (64 + lane < n) == (k < n), k);
                colv[buf][lane] = a[c][0];
                colv[buf][64 + lane] = a[c][1];
                if (lane == 0) pk[buf] = key;
            }
            __syncthreads();
            const int key = pk[buf];
            const bool piv0 = pos0 == key, piv1 = pos1 == key;
            const unsigned long long b0 = __ballot(piv0), b1 = __ballot(piv1);
            const int ph = b0 ? 0 : 1; // the half that holds the pivot row (uniform)
            const int pl = __ffsll(ph ? b1 : b0) - 1;
            if (pos0 == k)
                pos0 = key;
            else if (piv0)
                pos0 = k;
            if (pos1 == k)
                pos1 = key;
            else if (piv1)
                pos1 = k;
            const double pv = colv[buf][64 * ph + pl];
            if (pv == 0.0 && sinfo == 0) sinfo = k + 1;
            const double inv = 1.0 / pv, f0 = colv[buf][lane], f1 = colv[buf][64 + lane];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (wv == kb && j == c) continue;
                const double pj = gj_rdlane(ph ? a[j][1] : a[j][0], pl) * inv;
                a[j][0] = piv0 ? pj : a[j][0] - f0 * pj;
                a[j][1] = piv1 ? pj : a[j][1] - f1 * pj;
            }
            if (wv == kb) {
                a[c][0] = piv0 ? inv : -f0 * inv;
                a[c][1] = piv1 ? inv : -f1 * inv;
            }
        }
    }
    // column col of inv(PA) is column perm[col] of inv(A), perm[col] the
    // row at position col
    double *dst = Inv + (size_t)mat * istride;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int col = 16 * wv + j;
        const unsigned long long b0 = __ballot(pos0 == col), b1 = __ballot(pos1 == col);
        const int pc = b0 ? __ffsll(b0) - 1 : 64 + __ffsll(b1) - 1;
        if (col < n) {
            if (lane < n) dst[(size_t)pc * n + pos0] = a[j][0];
            if (64 + lane < n) dst[(size_t)pc * n + pos1] = a[j][1];
        }
    }
    if (threadIdx.x == 0) info[mat] = sinfo;
}

// Largest order the LDS Gauss-Jordan kernel takes on this device (0 if the
// dynamic-LDS limit cannot be raised).
static int gj_max_order()
{
    static int nmax = -1;
    if (nmax < 0) {
        int dev = 0, lds = 0;
        (void)hipGetDevice(&dev);
        // the most a kernel may opt in to (the default per-block limit is
        // 64 KB; this device's LDS holds 160 KB)
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess ||
            lds <= 65536) {
            (void)hipGetLastError();
            lds = 160 * 1024;
        }
        int want = lds - 8192; // room for the static scratch (4100 bytes)
        if (hipFuncSetAttribute((const void *)k_gj_inverse, hipFuncAttributeMaxDynamicSharedMemorySize,
                                want) != hipSuccess) {
            (void)hipGetLastError();
            want = 65536 - 4096;
        }
        int n = 0;
        while ((size_t)(n + 1) * (n + 1) * sizeof(double) <= (size_t)want) ++n;
        nmax = n;
    }
    return nmax;
}

// inv of `count` n x n matrices; LDS Gauss-Jordan when it fits, else global
// LU + inverse.  A is overwritten only on the fallback path.
static int batched_inverse(double *A, int n, int count, double *Inv, int *piv, int *info,
                           hipStream_t st)
{
    if (count <= 0 || n <= 0) return 0;
    if (n <= 64 && count <= 1024) { // a work-group per matrix (one wave per matrix beyond)
        if (n <= 32)
            hipLaunchKernelGGL((k_gj_inverse_q<32, GJQ_CW>), dim3(count), dim3(32 / GJQ_CW * 64), 0, st, A, n,
                               (size_t)n * n, Inv, (size_t)n * n, info);
        else
            hipLaunchKernelGGL((k_gj_inverse_q<64, GJQ_CW>), dim3(count), dim3(64 / GJQ_CW * 64), 0, st, A, n,
                               (size_t)n * n, Inv, (size_t)n * n, info);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    if (n <= 64) {
        const dim3 grid((count + 3) / 4);
        if (n <= 32)
            hipLaunchKernelGGL(k_gj_inverse_w<32>, grid, dim3(256), 0, st, A, n, (size_t)n * n, Inv,
                               (size_t)n * n, info, count);
        else
            hipLaunchKernelGGL(k_gj_inverse_w<64>, grid, dim3(256), 0, st, A, n, (size_t)n * n, Inv,
                               (size_t)n * n, info, count);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    if (n <= 128) {
        hipLaunchKernelGGL(k_gj_inverse_r, dim3(count), dim3(512), 0, st, A, n, (size_t)n * n, Inv,
                           (size_t)n * n, info);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    if (n <= gj_max_order()) {
        hipLaunchKernelGGL(k_gj_inverse, dim3(count), dim3(256), (size_t)n * n * sizeof(double), st,
                           A, n, (size_t)n * n, Inv, (size_t)n * n, info);
        PLR_HIP(hipGetLastError());
        return 0;
    }
    hipLaunchKernelGGL(k_lu_batched, dim3(count), dim3(256), 0, st, A, n, (size_t)n * n, piv, info);
    PLR_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_lu_inverse, dim3(count), dim3(n < 256 ? ((n + 63) & ~63) : 256), 0, st, A,
                       piv, n, (size_t)n * n, Inv, (size_t)n * n);
    PLR_HIP(hipGetLastError());
    return 0;
}

// Build K = [I, M_bot S_bot; M_top S_top, I] for every node of a level.
// M_s[k][c] = T[(off + c)*tstride + s*R + k] (output of the raw reduction
// over the level's own Ut columns).
__global__ void k_capf_build(const double *T, size_t tstride, int off, int R,
                             const double *S, double *K)
{
    const int node = blockIdx.x, n2 = 2 * R, ne = n2 * n2;
    double *k = K + (size_t)node * n2 * n2;
    const int top = 2 * node, bot = 2 * node + 1;
    // sixteen elements per thread and round, every load of a round issued
    // before the first is used (one memory wait per round)
    for (int e0 = 0; e0 < ne; e0 += 16 * (int)blockDim.x) {
        double tv[16], sv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = min(e0 + u * (int)blockDim.x + (int)threadIdx.x, ne - 1);
            const int i = e % n2, j = e / n2;
            // the off-diagonal blocks read T and S; the diagonal ones a valid
            // dummy (the first entries), multiplied by zero below
            const bool ur = i < R && j >= R, ll = i >= R && j < R;
            const size_t ti = ur ? (size_t)(off + j - R) * tstride + (size_t)bot * R + i
                                 : ll ? (size_t)(off + j) * tstride + (size_t)top * R + i - R : 0;
            const size_t si = ur ? (size_t)bot * R + j - R : ll ? (size_t)top * R + j : 0;
            tv[u] = T[ti];
            sv[u] = S[si];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = e0 + u * (int)blockDim.x + (int)threadIdx.x;
            if (e >= ne) continue;
            const int i = e % n2, j = e / n2;
            const bool od = (i < R) != (j < R);
            k[e] = (i == j ? 1.0 : 0.0) + (od ? tv[u] * sv[u] : 0.0);
        }
    }
}

// Capacitance step of the solve for one level (the kernel is in plr_cap.h):
// T is the raw reduction, W the negated expansion coefficients (layout of T).
static int launch_cap(const double *Kinv, int d, int R, const double *S, const double *T,
                      size_t tstride, double *W, int ncols, int trans, hipStream_t st)
{
    return plr_cap_launch<double, double, false>(Kinv, d, R, S, T, tstride, W, tstride, trans,
                                                 ncols, nullptr, 0, st);
}

// ---------------------------------------------------------------------------
// host side

static size_t col_off(const dplrh_plan *p, int d) { return p->N ? p->uoff[d] / p->N : 0; }

void plr_free_fact(dplrh_plan *p)
{
    ++p->ws_gen;
    if (p->f.gst) (void)hipStreamSynchronize(p->f.gst);
    void *bufs[] = {p->f.Dinv, p->f.Ut, p->f.Kinv, p->f.dinfo, p->f.work, p->f.G,
                    p->f.gC, p->f.gpart,
                    p->sw.T, p->sw.W, p->sw.part, p->sw.tmp, p->sw.cnt,
                    p->sw.uT0, p->sw.uW, p->sw.uA[0], p->sw.uA[1]};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (p->f.gev) (void)hipEventDestroy(p->f.gev);
    if (p->f.gdone) (void)hipEventDestroy(p->f.gdone);
    if (p->f.gst) (void)hipStreamDestroy(p->f.gst);
    memset(&p->f, 0, sizeof(p->f));
    memset(&p->sw, 0, sizeof(p->sw));
}

// Per-column sizes of the single-level reduction workspace (max over levels).
static void solve_ws_sizes(const dplrh_plan *p, size_t *t, size_t *part, size_t *cnt)
{
    LevelTab all;
    plr_level_tab(p, 0, &all);
    *t = *part = *cnt = 1;
    for (int d = 0; d < p->L; ++d) {
        const size_t nseg = (size_t)2 << d;
        if (nseg * all.R[d] > *t) *t = nseg * all.R[d];
        if (all.tps[d] > 1) {
            if (nseg * all.tps[d] * all.R[d] > *part) *part = nseg * all.tps[d] * all.R[d];
            if (nseg * all.nkg[d] > *cnt) *cnt = nseg * all.nkg[d];
        }
    }
}

static int solve_ws(dplrh_plan *p, int cols)
{
    plr_solve_ws &w = p->sw;
    if (cols <= w.cols) return 0;
    ++p->ws_gen;
    void *old[] = {w.T, w.W, w.part, w.cnt};
    for (void *b : old)
        if (b) (void)hipFree(b);
    w.T = w.W = w.part = nullptr;
    w.cnt = nullptr;
    w.cols = 0;
    solve_ws_sizes(p, &w.t_per_col, &w.part_per_col, &w.cnt_per_col);
    PLR_HIP(hipMalloc(&w.T, w.t_per_col * cols * sizeof(double)));
    PLR_HIP(hipMalloc(&w.W, w.t_per_col * cols * sizeof(double)));
    PLR_HIP(hipMalloc(&w.part, w.part_per_col * cols * sizeof(double)));
    PLR_HIP(hipMalloc(&w.cnt, w.cnt_per_col * cols * sizeof(int)));
    PLR_HIP(hipMemset(w.cnt, 0, w.cnt_per_col * cols * sizeof(int)));
    PLR_HIP(hipDeviceSynchronize());
    w.cols = cols;
    return 0;
}

static int solve_tmp(dplrh_plan *p, size_t elems)
{
    if (elems <= p->sw.tmp_cap) return 0;
    ++p->ws_gen;
    if (p->sw.tmp) (void)hipFree(p->sw.tmp);
    p->sw.tmp = nullptr;
    p->sw.tmp_cap = 0;
    PLR_HIP(hipMalloc(&p->sw.tmp, elems * sizeof(double)));
    p->sw.tmp_cap = elems;
    return 0;
}

// Post-order position of a failure: a node is visited after its last leaf
// and after every deeper node that ends at the same leaf.
struct fail_key
{
    long last_leaf;
    int height;
    int info;
};

static bool key_less(const fail_key &a, const fail_key &b)
{
    return a.last_leaf != b.last_leaf ? a.last_leaf < b.last_leaf : a.height < b.height;
}

static int read_info(int count, const int *dinfo, std::vector<int> &h, hipStream_t st)
{
    h.resize(count > 0 ? count : 1);
    PLR_HIP(hipStreamSynchronize(st));
    if (count > 0) PLR_HIP(hipMemcpy(h.data(), dinfo, count * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

// FP64 factorization; *info as dgetrf reports it for the first singular
// matrix in the reference's post-order (0 if none).
extern "C" int dplrh_factor(dplrh_plan *p, int *info, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    *info = 0;
    plr_fact &f = p->f;
    // G work of an earlier call that stopped on an error may still read Ut
    if (f.gst) PLR_HIP(hipStreamSynchronize(f.gst));
    const int N = p->N, L = p->L, b = p->b;
    const int nleaf = b ? N / b : 0;
    if (N == 0) {
        f.done = 1;
        return 0;
    }
    if (!f.Dinv) {
        size_t kk = 0, work = (size_t)N * b, piv = (size_t)nleaf * b;
        int maxnodes = nleaf;
        for (int d = 0; d < L; ++d) {
            f.koff[d] = kk;
            const size_t n2 = 2 * (size_t)p->R[d];
            kk += ((size_t)1 << d) * n2 * n2;
            if (((size_t)1 << d) * n2 * n2 > work) work = ((size_t)1 << d) * n2 * n2;
            if (((size_t)1 << d) * n2 > piv) piv = ((size_t)1 << d) * n2;
            if ((1 << d) > maxnodes) maxnodes = 1 << d;
        }
        f.koff[L] = kk;
        PLR_HIP(hipMalloc(&f.Dinv, (size_t)N * b * sizeof(double)));
        PLR_HIP(hipMalloc(&f.Ut, (p->uoff[L] ? p->uoff[L] : 1) * sizeof(double)));
        PLR_HIP(hipMalloc(&f.Kinv, (kk ? kk : 1) * sizeof(double)));
        PLR_HIP(hipMalloc(&f.work, (work + piv / 2 + 1) * sizeof(double) + piv * sizeof(int)));
        // LU status of every leaf, then of every node level by level (level d
        // at nleaf + 2^d - 1), read once at the end
        (void)maxnodes;
        PLR_HIP(hipMalloc(&f.dinfo, ((size_t)nleaf + ((size_t)1 << L)) * sizeof(int)));
    }
    f.done = 0;
    ++p->ws_gen; // factors are rewritten: graphs replaying solves stay valid, but be safe
    double *work = f.work;
    // pivots live behind the largest LU scratch
    size_t worksz = (size_t)N * b;
    for (int d = 0; d < L; ++d) {
        const size_t n2 = 2 * (size_t)p->R[d];
        if (((size_t)1 << d) * n2 * n2 > worksz) worksz = ((size_t)1 << d) * n2 * n2;
    }
    int *piv = (int *)(work + worksz);

    std::vector<fail_key> fails;
    std::vector<int> h;

    // 1. leaves
    PLR_HIP(hipMemcpyAsync(work, p->D, (size_t)N * b * sizeof(double), hipMemcpyDeviceToDevice, st));
    {
        int rc = batched_inverse(work, b, nleaf, f.Dinv, piv, f.dinfo, st);
        if (rc) return rc;
    }

    // 2. Ut = Dinv U (all levels' columns)
    const int ctot = (int)col_off(p, L);
    if (ctot > 0) {
        LevelTab t0;
        plr_level_tab0(p, &t0);
        int rc = 1;
        if (plr_mm_ok(p) && !p->force_valu)
            rc = plr_mm_expand_lvl(p, -1, nullptr, nullptr, 0, f.Dinv, p->U, N, 1.0, 0.0, f.Ut, N,
                                   ctot, st);
        if (rc == 1 && plr_mm_ok(p) && !p->force_valu)
            rc = plr_mm_expand1(p, -1, nullptr, nullptr, 0, f.Dinv, 0, p->U, N, 1.0, 0.0, f.Ut, N,
                                ctot, st);
        else if (rc == 1)
            rc = plr_launch_expand(t0, f.Dinv, p->U, N, 1.0, 0.0, 0, nullptr, 0, f.Ut, N, ctot, st);
        if (rc) return rc;
    }

    // 3. levels, bottom-up; the G blocks of the reassociated upper solve
    // beside them on their own stream when the shape allows
    const int gd0 = plr_up_begin(p);
    if (gd0 < 0) return gd0;
    for (int d = L - 1; d >= 0; --d) {
        const int R = p->R[d];
        if (R == 0) continue;
        const int off = (int)col_off(p, d);
        const int ncol = off + R;
        // the ancestor update through the third-form expansion takes the
        // capacitance output rank-major over whole 64-column groups
        const bool fx = plr_mm_ok(p) && !p->force_valu && N % 256 == 0;
        const int wtc = (off + 63) / 64 * 64;
        int rc = solve_ws(p, fx && wtc > ncol ? wtc : ncol);
        if (rc) return rc;
        const double *Ud = f.Ut + p->uoff[d], *Vd = p->V + p->uoff[d], *Sd = p->S + p->soff[d];
        LevelTab t1;
        plr_level_tab1(p, d, Vd, Ud, Sd, &t1);
        const bool mm = plr_mm_ok(p) && !p->force_valu;
        if (mm)
            rc = plr_mm_reduce1(p, d, Vd, f.Ut, N, ncol, p->sw.T, p->sw.t_per_col, st);
        else
            rc = plr_launch_reduce(t1, f.Ut, N, 1.0, 0, 1, ncol, p->sw.part, p->sw.part_per_col,
                                   p->sw.cnt, p->sw.cnt_per_col, p->sw.T, p->sw.t_per_col, st);
        if (rc) return rc;
        const int nodes = 1 << d, n2 = 2 * R;
        hipLaunchKernelGGL(k_capf_build, dim3(nodes), dim3(256), 0, st, p->sw.T, p->sw.t_per_col,
                           off, R, Sd, work);
        PLR_HIP(hipGetLastError());
        // G of level d beside this level's inverses (few work-groups)
        if (gd0 > 0 && d >= 1 && d < gd0 && (rc = plr_up_level(p, d, st))) return rc;
        if ((rc = batched_inverse(work, n2, nodes, f.Kinv + f.koff[d], piv,
                                  f.dinfo + nleaf + (nodes - 1), st)))
            return rc;
        if (off > 0 && fx) {
            // W rank-major: row s*R + k, column c at W[(s*R + k) * wtc + c]
            if ((rc = plr_cap_launch<double, double, false>(f.Kinv + f.koff[d], d, R, Sd, p->sw.T,
                                                            p->sw.t_per_col, p->sw.W, 1, 0, off,
                                                            nullptr, 0, st, (size_t)wtc)))
                return rc;
            if ((rc = plr_mm_expand_lvl(p, d, Ud, p->sw.W, wtc, nullptr, nullptr, 0, 1.0, 1.0, f.Ut,
                                        N, off, st)))
                return rc < 0 ? rc : -3;
        } else if (off > 0) {
            if ((rc = launch_cap(f.Kinv + f.koff[d], d, R, Sd, p->sw.T, p->sw.t_per_col, p->sw.W, off,
                                 0, st)))
                return rc;
            if (mm)
                rc = plr_mm_expand1(p, d, Ud, p->sw.W, p->sw.t_per_col, nullptr, 0, nullptr, N,
                                    1.0, 1.0, f.Ut, N, off, st);
            else
                rc = plr_launch_expand(t1, nullptr, nullptr, N, 1.0, 1.0, 0, p->sw.W,
                                       p->sw.t_per_col, f.Ut, N, off, st);
            if (rc) return rc;
        }
    }
    if (gd0 > 0) {
        const int rc = plr_up_end(p, st);
        if (rc) return rc;
    }
    // every LU status at once: leaves, then the nodes of each level with a rank
    if (read_info(nleaf + (1 << L) - 1, f.dinfo, h, st)) return -10;
    for (int l = 0; l < nleaf; ++l)
        if (h[l]) fails.push_back({l, 0, h[l]});
    for (int d = 0; d < L; ++d) {
        if (p->R[d] == 0) continue;
        for (int n = 0; n < (1 << d); ++n) {
            const int v = h[nleaf + (1 << d) - 1 + n];
            if (v) fails.push_back({((long)(n + 1) << (L - d)) - 1, L - d, v});
        }
    }
    // the cross-level blocks of the reassociated upper solve
    if (gd0 > 0) {
        f.up_d0 = fails.empty() ? gd0 : 0;
    } else if (fails.empty()) {
        const int rc = plr_up_factor(p, st);
        if (rc) return rc;
    }
    PLR_HIP(hipStreamSynchronize(st));
    if (!fails.empty()) {
        fail_key best = fails[0];
        for (const fail_key &k : fails)
            if (key_less(k, best)) best = k;
        *info = best.info;
        f.info = best.info;
        return 0;
    }
    f.info = 0;
    f.done = 1;
    return 0;
}

// X := op(A)^{-1} B for n device columns (ldb, ldx >= N), after dplrh_factor.
int plr_solve_fast(dplrh_plan *p, const double *Bm, int ldb, double *Xm, int ldx, int ncols,
                   hipStream_t st);
int plr_solve_tail(dplrh_plan *p, const double *Bm, int ldb, double *Xm, int ldx, int ncols,
                   hipStream_t st);

static int solve_body(dplrh_plan *p, int trans, int n, const double *dB, int ldb, double *dX,
                      int ldx, hipStream_t st)
{
    const int N = p->N, L = p->L;
    if (N == 0 || n == 0) return 0;
    const int chunk = n < 1024 ? n : 1024;
    int rc = solve_ws(p, chunk);
    if (rc) return rc;
    LevelTab t0;
    plr_level_tab0(p, &t0);
    const bool mm = plr_mm_ok(p) && !p->force_valu && ldb % 2 == 0 && ldx % 2 == 0 && n >= 2;
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int nc = n - c0 < chunk ? n - c0 : chunk;
        const double *B = dB + (size_t)c0 * ldb;
        double *X = dX + (size_t)c0 * ldx;
        double *Z = X;
        int ldz = ldx;
        int dstop = L; // levels [dstop, L) and the leaves done by the tail kernel
        // the fused path reads B and X with scalar loads: any leading
        // dimension and a single column are fine
        if (!trans && plr_mm_ok(p) && !p->force_valu) {
            rc = plr_solve_fast(p, B, ldb, X, ldx, nc, st);
            if (rc < 0) return rc;
            if (rc == 0) continue;
            rc = 0;
        }
        if (!trans) {
            const int d0 = mm ? plr_solve_tail(p, B, ldb, X, ldx, nc, st) : -1;
            if (d0 < -1) return d0; // a HIP error (-1: shape not covered)
            if (d0 >= 0)
                dstop = d0;
            else if (mm)
                rc = plr_mm_expand1(p, -1, nullptr, nullptr, 0, p->f.Dinv, 0, B, ldb, 1.0, 0.0, X,
                                    ldx, nc, st);
            else
                rc = plr_launch_expand(t0, p->f.Dinv, B, ldb, 1.0, 0.0, 0, nullptr, 0, X, ldx, nc,
                                       st);
            if (rc) return rc;
        } else {
            if ((rc = solve_tmp(p, (size_t)N * nc))) return rc;
            Z = p->sw.tmp;
            ldz = N;
            PLR_HIP(hipMemcpy2DAsync(Z, (size_t)N * sizeof(double), B, (size_t)ldb * sizeof(double),
                                     (size_t)N * sizeof(double), nc, hipMemcpyDeviceToDevice, st));
        }
        for (int k = 0; k < L; ++k) {
            const int d = trans ? k : L - 1 - k;
            const int R = p->R[d];
            if (R == 0 || d >= dstop) continue;
            const double *Ud = p->f.Ut + p->uoff[d], *Vd = p->V + p->uoff[d];
            const double *Sd = p->S + p->soff[d];
            LevelTab t1;
            plr_level_tab1(p, d, trans ? Ud : Vd, trans ? Vd : Ud, Sd, &t1);
            if (mm)
                rc = plr_mm_reduce1(p, d, trans ? Ud : Vd, Z, ldz, nc, p->sw.T, p->sw.t_per_col, st);
            else
                rc = plr_launch_reduce(t1, Z, ldz, 1.0, 0, 1, nc, p->sw.part, p->sw.part_per_col,
                                       p->sw.cnt, p->sw.cnt_per_col, p->sw.T, p->sw.t_per_col, st);
            if (rc) return rc;
            if ((rc = launch_cap(p->f.Kinv + p->f.koff[d], d, R, Sd, p->sw.T, p->sw.t_per_col,
                                 p->sw.W, nc, trans, st)))
                return rc;
            if (mm)
                rc = plr_mm_expand1(p, d, trans ? Vd : Ud, p->sw.W, p->sw.t_per_col, nullptr, 0,
                                    nullptr, ldz, 1.0, 1.0, Z, ldz, nc, st);
            else
                rc = plr_launch_expand(t1, nullptr, nullptr, ldz, 1.0, 1.0, 0, p->sw.W,
                                       p->sw.t_per_col, Z, ldz, nc, st);
            if (rc) return rc;
        }
        if (trans) {
            if (mm)
                rc = plr_mm_expand1(p, -1, nullptr, nullptr, 0, p->f.Dinv, 1, Z, ldz, 1.0, 0.0, X,
                                    ldx, nc, st);
            else
                rc = plr_launch_expand(t0, p->f.Dinv, Z, ldz, 1.0, 0.0, 1, nullptr, 0, X, ldx, nc,
                                       st);
            if (rc) return rc;
        }
    }
    return 0;
}

int dplrh_lp_lu_inverse(double *K, int n, int count, double *Kinv, int *info, hipStream_t st)
{
    *info = 0;
    if (count <= 0 || n <= 0) return 0;
    int *piv = nullptr, *dinfo = nullptr;
    PLR_HIP(hipMalloc(&piv, (size_t)count * n * sizeof(int)));
    PLR_HIP(hipMalloc(&dinfo, (size_t)count * sizeof(int)));
    int rc0 = batched_inverse(K, n, count, Kinv, piv, dinfo, st);
    if (rc0) return rc0;
    std::vector<int> h;
    int rc = read_info(count, dinfo, h, st);
    (void)hipFree(piv);
    (void)hipFree(dinfo);
    if (rc) return rc;
    for (int i = 0; i < count; ++i)
        if (h[i]) {
            *info = h[i];
            break;
        }
    return 0;
}

extern "C" int dplrh_solve(dplrh_plan *p, int trans, int n, const double *dB, int ldb,
                           double *dX, int ldx, void *stream)
{
    if (!p->f.done) {
        plr_set_error("solve before a successful factorization");
        return -4;
    }
    if (p->N == 0 || n == 0) return 0;
    const unsigned long long key[10] = {3, (unsigned long long)trans, (unsigned long long)n,
                                        (unsigned long long)dB, (unsigned long long)ldb,
                                        (unsigned long long)dX, (unsigned long long)ldx};
    return plr_run_cached(p, key, (hipStream_t)stream, [&](hipStream_t s) {
        return solve_body(p, trans, n, dB, ldb, dX, ldx, s);
    });
}

// ---------------------------------------------------------------------------
// Single-level building blocks with MFMA / general-shape dispatch.

// T[c][s*R_d + k] = P_d|_s^T X|_s for ncols columns (raw, FP64).
int plr_lvl_reduce(dplrh_plan *p, int d, const double *P, const double *X, int ldx, int ncols,
                   double *T, size_t tstride, hipStream_t st)
{
    if (ncols <= 0 || p->R[d] == 0) return 0;
    if (plr_mm_ok(p) && !p->force_valu && ldx % 2 == 0)
        return plr_mm_reduce1(p, d, P, X, ldx, ncols, T, tstride, st);
    int rc = solve_ws(p, ncols);
    if (rc) return rc;
    LevelTab t1;
    plr_level_tab1(p, d, P, nullptr, nullptr, &t1);
    return plr_launch_reduce(t1, X, ldx, 1.0, 0, 1, ncols, p->sw.part, p->sw.part_per_col,
                             p->sw.cnt, p->sw.cnt_per_col, T, tstride, st);
}

// Y := beta*Y + Q_d W (per segment of level d), or with d < 0 the leaf term
// Y := alpha*op(D) X + beta*Y.
int plr_lvl_expand(dplrh_plan *p, int d, const double *Q, const double *W, size_t wstride,
                   const double *D, int trans, const double *X, int ldx, double alpha,
                   double beta, double *Y, int ldy, int ncols, hipStream_t st)
{
    if (ncols <= 0) return 0;
    if (plr_mm_ok(p) && !p->force_valu && (d >= 0 || ldx % 2 == 0))
        return plr_mm_expand1(p, d, Q, W, wstride, D, trans, X, ldx, alpha, beta, Y, ldy, ncols,
                              st);
    LevelTab t;
    if (d >= 0)
        plr_level_tab1(p, d, nullptr, Q, nullptr, &t);
    else
        plr_level_tab0(p, &t);
    return plr_launch_expand(t, D, X, ldx, alpha, beta, trans, W, wstride, Y, ldy, ncols, st);
}

// W = capacitance coefficients for one level (see k_cap_apply).
int plr_lvl_cap(dplrh_plan *p, int d, const double *T, size_t tstride, double *W, int ncols,
                int trans, hipStream_t st)
{
    const int R = p->R[d];
    if (ncols <= 0 || R == 0) return 0;
    return launch_cap(p->f.Kinv + p->f.koff[d], d, R, p->S + p->soff[d], T, tstride, W, ncols,
                      trans, st);
}

int plr_solve_reserve(dplrh_plan *p, int cols) { return solve_ws(p, cols); }
