// Compression of a dense host matrix into a PLR tree: the reference's
// dplrd2gep (src/dplrd2p.c:187) with the same tree shaping, the same split
// of eps over the levels (eps/2 at the last level, eps/4 above, per level of
// recursion) and the same rank rule (the number of leading singular values
// >= the block's share of eps, src/dplrd2p.c:126-127).  The reference takes
// a full SVD of every block (dgesvd_, src/dplrd2p.c:81,110); here every
// off-diagonal block of one level is gathered into a strided batch and its
// leading singular triplets come from a truncated SVD written for the GPU:
//   1. range:   Y = B W (W Gaussian, len x s), Q = qr(Y), then one power
//               step Q = qr(B qr(B^T Q))  (Householder QR, one work-group
//               per block)
//   2. project: M = B^T Q (len x s), so that B ~ Q M^T
//   3. SVD:     one-sided Jacobi (Hestenes) on M: M J has orthogonal
//               columns, J orthogonal (s x s); sigma_k = |M J e_k|,
//               v_k = M J e_k / sigma_k, u_k = Q J e_k
// The products are plain batched FP64 GEMMs (rocBLAS).  A block's sample is
// large enough when its smallest computed singular value is below the
// threshold and either below 1e-4 times it or at the rounding floor (1e3 eps
// times the largest): then the kept triplets are accurate to far below the
// threshold, or to the rounding level of any backward-stable SVD
// (Rayleigh-Ritz errors are second order in the first omitted singular
// value).  Otherwise the level is sampled again with twice the columns, up
// to s = len, where the decomposition is exact.  Singular vectors may
// differ in sign from LAPACK's (u_k v_k^T does not).
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dplr.h"
#include "plr_common.h"

// out[blk][j*len + i] = A(row0(blk) + i, col0(blk) + j); block 2n is node n's
// ne block (rows of segment 2n, columns of segment 2n+1), 2n+1 its sw block.
__global__ void k_d2p_gather(const double *__restrict__ A, size_t lda, int len, int nblk,
                             double *__restrict__ out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)len * len;
    if (e >= per * nblk) return;
    const int blk = (int)(e / per);
    const size_t r = e % per;
    const int i = (int)(r % len), j = (int)(r / len);
    const size_t row = (size_t)blk * len + i, col = (size_t)(blk ^ 1) * len + j;
    out[e] = A[col * lda + row];
}

static double d2p_eps(double eps, int d, int nlvl)
{
    for (int k = 0; k < d; ++k) eps /= 4.0;
    return (nlvl - d == 1) ? eps / 2.0 : eps / 4.0;
}

// The reference's shaping: an H node with two rank-0 low-rank blocks per
// level, dense leaves at depth nlvl.
static void d2p_shape(DPlrMtx *B, int m, int nlvl)
{
    if (nlvl == 0) {
        if (B->type != PLR_DMTX || B->size != m) {
            dplrclr(B);
            dplrinid(m, B);
        }
        return;
    }
    if (B->type != PLR_HMTX || B->size != m) {
        dplrclr(B);
        dplrinih(m, B);
        dplrinilr(m / 2, B->dh.ne, 0);
        dplrinilr(m / 2, B->dh.sw, 0);
    }
    d2p_shape(B->dh.nw, m / 2, nlvl - 1);
    d2p_shape(B->dh.se, m / 2, nlvl - 1);
}

static DPlrMtx *d2p_node(DPlrMtx *B, int d, int n)
{
    for (int k = d - 1; k >= 0; --k) B = ((n >> k) & 1) ? B->dh.se : B->dh.nw;
    return B;
}

// ---- truncated SVD kernels -------------------------------------------------

__device__ __forceinline__ unsigned long long d2p_mix(unsigned long long x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

// N(0,1) samples (Box-Muller on two 53-bit uniforms in (0, 1])
__global__ void k_d2p_normal(double *__restrict__ out, size_t n, unsigned long long seed)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long a = d2p_mix(seed ^ (2 * i)), b = d2p_mix(seed ^ (2 * i + 1));
    const double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
    out[i] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

// Sums of NV values over the 256 threads of the work-group (every thread
// gets the totals); lds holds 4 x NV doubles.
template <int NV>
__device__ __forceinline__ void d2p_wg_sum(double (&v)[NV], double *lds)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        v[k] = x;
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[w * NV + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k)
        v[k] = (lds[k] + lds[NV + k]) + (lds[2 * NV + k] + lds[3 * NV + k]);
    __syncthreads();
}

// Y (len x s, column-major, ld len) := the Q of its thin Householder QR, for
// every block of the batch (one work-group of 256 threads each).  Row i is
// always handled by thread i % 256, so a thread only reads rows it wrote
// itself except through the work-group sums.
__global__ __launch_bounds__(256) void k_d2p_qr(double *__restrict__ Yall, int len, int s,
                                                size_t stride, double *__restrict__ tauall)
{
    __shared__ double lds[4 * 8];
    double *Y = Yall + (size_t)blockIdx.x * stride;
    double *tau = tauall + (size_t)blockIdx.x * s;
    const int tid = threadIdx.x;
    // factorization: H_j = I - tau_j v_j v_j^T, v_j = [1; Y(j+1:, j)]
    for (int j = 0; j < s; ++j) {
        double *y = Y + (size_t)j * len;
        double nv[1] = {0.0};
        for (int i = tid; i < len; i += 256)
            if (i > j) nv[0] += y[i] * y[i];
        d2p_wg_sum<1>(nv, lds);
        const double below = nv[0], alpha = y[j];
        double t = 0.0, beta = alpha, scale = 0.0;
        if (below > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + below), alpha);
            t = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        __syncthreads(); // every thread has read alpha
        for (int i = tid; i < len; i += 256) {
            if (i > j) y[i] *= scale;
            if (i == j) y[i] = beta;
        }
        if (tid == 0) tau[j] = t;
        if (t == 0.0) continue;
        for (int c0 = j + 1; c0 < s; c0 += 8) {
            const int nc = min(8, s - c0);
            double w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = tid; i < len; i += 256) {
                if (i < j) continue;
                const double vi = i == j ? 1.0 : y[i];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k < nc) w[k] += vi * Y[(size_t)(c0 + k) * len + i];
            }
            d2p_wg_sum<8>(w, lds);
            for (int i = tid; i < len; i += 256) {
                if (i < j) continue;
                const double vi = i == j ? 1.0 : y[i];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k < nc) Y[(size_t)(c0 + k) * len + i] -= t * vi * w[k];
            }
        }
    }
    __syncthreads(); // tau visible to every thread
    // Q = H_0 ... H_{s-1} [I; 0], formed in place from the last column back
    for (int j = s - 1; j >= 0; --j) {
        double *y = Y + (size_t)j * len;
        const double t = tau[j];
        for (int c0 = j + 1; c0 < s && t != 0.0; c0 += 8) {
            const int nc = min(8, s - c0);
            double w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = tid; i < len; i += 256) {
                if (i < j) continue;
                const double vi = i == j ? 1.0 : y[i];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k < nc) w[k] += vi * Y[(size_t)(c0 + k) * len + i];
            }
            d2p_wg_sum<8>(w, lds);
            for (int i = tid; i < len; i += 256) {
                if (i < j) continue;
                const double vi = i == j ? 1.0 : y[i];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k < nc) Y[(size_t)(c0 + k) * len + i] -= t * vi * w[k];
            }
        }
        for (int i = tid; i < len; i += 256) y[i] = i > j ? -t * y[i] : (i == j ? 1.0 - t : 0.0);
    }
}

// One-sided Jacobi: M (len x s) := M J with mutually orthogonal columns, J
// (s x s) the accumulated rotations, for every block of the batch.  Column
// pairs follow the circle ordering (s - 1 rounds of disjoint pairs per
// sweep, a dummy column when s is odd); the four waves take the pairs of a
// round, each pair's sums over 64 lanes.  Sweeps stop when no pair has
// |m_p^T m_q| > tol |m_p| |m_q|, or after maxsweep.
__global__ __launch_bounds__(256) void k_d2p_jacobi(double *__restrict__ Mall, int len, int s,
                                                    size_t mstride, double *__restrict__ Jall,
                                                    double tol, int maxsweep)
{
    __shared__ unsigned long long offbits;
    double *M = Mall + (size_t)blockIdx.x * mstride;
    double *J = Jall + (size_t)blockIdx.x * s * s;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int e = tid; e < s * s; e += 256) J[e] = (e % s == e / s) ? 1.0 : 0.0;
    const int n = s + (s & 1);
    for (int sweep = 0; sweep < maxsweep; ++sweep) {
        if (tid == 0) offbits = 0ull;
        __syncthreads();
        double off = 0.0;
        for (int r = 0; r < n - 1; ++r) {
            for (int k = w; k < n / 2; k += 4) {
                const int p = k == 0 ? n - 1 : (r + k) % (n - 1);
                const int q = k == 0 ? r : (r - k + (n - 1)) % (n - 1);
                if (p >= s || q >= s) continue;
                double *mp = M + (size_t)p * len, *mq = M + (size_t)q * len;
                double a = 0.0, b = 0.0, g = 0.0;
                for (int i = lane; i < len; i += 64) {
                    const double x = mp[i], y = mq[i];
                    a += x * x;
                    b += y * y;
                    g += x * y;
                }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    a += __shfl_xor(a, o);
                    b += __shfl_xor(b, o);
                    g += __shfl_xor(g, o);
                }
                if (!(a > 0.0 && b > 0.0)) continue;
                const double rel = fabs(g) / sqrt(a * b);
                if (!(rel > tol)) continue;
                off = fmax(off, rel);
                // zero m_p'^T m_q' with m_p' = c m_p - sn m_q, m_q' = sn m_p + c m_q
                const double zeta = (b - a) / (2.0 * g);
                const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                for (int i = lane; i < len; i += 64) {
                    const double x = mp[i], y = mq[i];
                    mp[i] = c * x - sn * y;
                    mq[i] = sn * x + c * y;
                }
                double *jp = J + (size_t)p * s, *jq = J + (size_t)q * s;
                for (int i = lane; i < s; i += 64) {
                    const double x = jp[i], y = jq[i];
                    jp[i] = c * x - sn * y;
                    jq[i] = sn * x + c * y;
                }
            }
            __syncthreads();
        }
        if (lane == 0) atomicMax(&offbits, (unsigned long long)__double_as_longlong(off));
        __syncthreads();
        const double total = __longlong_as_double((long long)offbits);
        __syncthreads();
        if (!(total > tol)) break;
    }
}

#define D2P_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            fprintf(stderr, "dplrd2gep: %s: %s\n", #call, hipGetErrorString(e_));                  \
            goto out;                                                                              \
        }                                                                                          \
    } while (0)
#define D2P_BLAS(call)                                                                             \
    do {                                                                                           \
        rocblas_status s_ = (call);                                                                \
        if (s_ != rocblas_status_success) {                                                        \
            fprintf(stderr, "dplrd2gep: %s: status %d\n", #call, (int)s_);                         \
            goto out;                                                                              \
        }                                                                                          \
    } while (0)

// Truncated SVDs of nblk blocks (len x len, column-major, stride len^2 in dB)
// with s sampled columns: on return hM holds M J (len x s per block) and hU
// holds Q J.  Buffers: dY, dZ, dM, dU of nblk * len * s doubles, dW of
// len * s, dJ of nblk * s * s, dtau of nblk * s.
static constexpr double kEps = 2.220446049250313e-16;

struct D2pBufs
{
    double *dY, *dZ, *dM, *dU, *dW, *dJ, *dtau;
};

static int d2p_tsvd(rocblas_handle h, const double *dB, int len, int nblk, int s, int level,
                    D2pBufs &b, std::vector<double> &hM, std::vector<double> &hU)
{
    const double one = 1.0, zero = 0.0;
    const size_t per = (size_t)len * len, ls = (size_t)len * s, tot = ls * nblk;
    const double tol = sqrt((double)len) * kEps;
    hipLaunchKernelGGL(k_d2p_normal, dim3((unsigned)((ls + 255) / 256)), dim3(256), 0, 0, b.dW, ls,
                       0x5eed0000ull + (unsigned long long)level * 131 + s);
    D2P_HIP(hipGetLastError());
    // Y = B W, Q = qr(Y)
    D2P_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, len,
                                           s, len, &one, dB, len, per, b.dW, len, 0, &zero, b.dY,
                                           len, ls, nblk));
    hipLaunchKernelGGL(k_d2p_qr, dim3(nblk), dim3(256), 0, 0, b.dY, len, s, ls, b.dtau);
    // power step: Z = qr(B^T Q), Y = qr(B Z)
    D2P_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none,
                                           len, s, len, &one, dB, len, per, b.dY, len, ls, &zero,
                                           b.dZ, len, ls, nblk));
    hipLaunchKernelGGL(k_d2p_qr, dim3(nblk), dim3(256), 0, 0, b.dZ, len, s, ls, b.dtau);
    D2P_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, len,
                                           s, len, &one, dB, len, per, b.dZ, len, ls, &zero, b.dY,
                                           len, ls, nblk));
    hipLaunchKernelGGL(k_d2p_qr, dim3(nblk), dim3(256), 0, 0, b.dY, len, s, ls, b.dtau);
    // M = B^T Q, then Jacobi: M J, J
    D2P_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none,
                                           len, s, len, &one, dB, len, per, b.dY, len, ls, &zero,
                                           b.dM, len, ls, nblk));
    hipLaunchKernelGGL(k_d2p_jacobi, dim3(nblk), dim3(256), 0, 0, b.dM, len, s, ls, b.dJ, tol, 60);
    D2P_HIP(hipGetLastError());
    // U = Q J
    D2P_BLAS(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, len,
                                           s, s, &one, b.dY, len, ls, b.dJ, s, (size_t)s * s, &zero,
                                           b.dU, len, ls, nblk));
    hM.resize(tot);
    hU.resize(tot);
    D2P_HIP(hipMemcpy(hM.data(), b.dM, tot * sizeof(double), hipMemcpyDeviceToHost));
    D2P_HIP(hipMemcpy(hU.data(), b.dU, tot * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
out:
    return -1;
}

extern "C" void dplrd2gep(int m, double *A, int lda, DPlrMtx *B, double eps, int nlvl, int *info)
{
    *info = -1;
    if (m < 0) {
        fprintf(stderr, "%s: invalid matrix size %d\n", __func__, m);
        return;
    }
    if (lda < m) {
        fprintf(stderr, "%s: invalid leading dimension\n", __func__);
        return;
    }
    if (eps < 0.0) {
        fprintf(stderr, "%s: eps must be nonnegative\n", __func__);
        return;
    }
    if (nlvl < 0 || nlvl > 30 || m % (1 << nlvl) != 0) {
        fprintf(stderr, "%s: m not divisible by 2**nlvl\n", __func__);
        return;
    }
    d2p_shape(B, m, nlvl);
    // leaves: dense copies of the diagonal blocks (host)
    {
        const int b = m >> nlvl;
        for (int l = 0; l < (1 << nlvl); ++l) {
            DPlrMtx *leaf = d2p_node(B, nlvl, l);
            for (int j = 0; j < b; ++j)
                memcpy(leaf->dd.d + (size_t)j * b, A + ((size_t)l * b + j) * lda + (size_t)l * b,
                       (size_t)b * sizeof(double));
        }
    }
    if (nlvl == 0 || m == 0) {
        *info = 0;
        return;
    }
    {
        double *dA = nullptr, *dB = nullptr;
        D2pBufs bf = {};
        size_t bcap = 0; // doubles per buffer held
        rocblas_handle h = nullptr;
        int failed = 1;
        const size_t half = (size_t)m * m / 2; // every level's blocks cover half of A
        std::vector<double> hM, hU, sg;
        std::vector<int> ord;
        if (rocblas_create_handle(&h) != rocblas_status_success) {
            fprintf(stderr, "%s: rocblas_create_handle failed\n", __func__);
            return;
        }
        D2P_HIP(hipMalloc(&dA, (size_t)m * m * sizeof(double)));
        D2P_HIP(hipMalloc(&dB, half * sizeof(double)));
        D2P_HIP(hipMemcpy2D(dA, (size_t)m * sizeof(double), A, (size_t)lda * sizeof(double),
                            (size_t)m * sizeof(double), m, hipMemcpyHostToDevice));
        for (int d = 0; d < nlvl; ++d) {
            const int len = m >> (d + 1), nblk = 2 << d;
            const size_t per = (size_t)len * len, tot = per * nblk;
            const double e2 = d2p_eps(eps, d, nlvl);
            hipLaunchKernelGGL(k_d2p_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0,
                               dA, (size_t)m, len, nblk, dB);
            D2P_HIP(hipGetLastError());
            for (int s = std::min(len, 64);; s = std::min(len, 2 * s)) {
                const size_t need = (size_t)nblk * len * s;
                if (need > bcap) {
                    double **bufs[] = {&bf.dY, &bf.dZ, &bf.dM, &bf.dU, &bf.dW, &bf.dJ, &bf.dtau};
                    for (double **p : bufs)
                        if (*p) (void)hipFree(*p);
                    bf = D2pBufs{};
                    D2P_HIP(hipMalloc(&bf.dY, need * sizeof(double)));
                    D2P_HIP(hipMalloc(&bf.dZ, need * sizeof(double)));
                    D2P_HIP(hipMalloc(&bf.dM, need * sizeof(double)));
                    D2P_HIP(hipMalloc(&bf.dU, need * sizeof(double)));
                    D2P_HIP(hipMalloc(&bf.dW, need * sizeof(double)));
                    D2P_HIP(hipMalloc(&bf.dJ, need * sizeof(double))); // nblk s^2 <= need
                    D2P_HIP(hipMalloc(&bf.dtau, need * sizeof(double)));
                    bcap = need;
                }
                if (d2p_tsvd(h, dB, len, nblk, s, d, bf, hM, hU)) goto out;
                // singular values, and whether every block's sample reached below the threshold
                const size_t ls = (size_t)len * s;
                sg.assign((size_t)nblk * s, 0.0);
                bool enough = true;
                for (int q = 0; q < nblk; ++q) {
                    double smin = INFINITY, smax = 0.0;
                    for (int k = 0; k < s; ++k) {
                        const double *mk = hM.data() + q * ls + (size_t)k * len;
                        double acc = 0.0;
                        for (int i = 0; i < len; ++i) acc += mk[i] * mk[i];
                        sg[(size_t)q * s + k] = sqrt(acc);
                        smin = std::min(smin, sg[(size_t)q * s + k]);
                        smax = std::max(smax, sg[(size_t)q * s + k]);
                    }
                    if (!(smin < e2 && (smin < 1e-4 * e2 || smin < 1e3 * kEps * smax)))
                        enough = false;
                }
                if (!enough && s < len) continue;
                for (int q = 0; q < nblk; ++q) {
                    const double *sq = sg.data() + (size_t)q * s;
                    ord.resize(s);
                    for (int k = 0; k < s; ++k) ord[k] = k;
                    std::stable_sort(ord.begin(), ord.end(),
                                     [&](int x, int y) { return sq[x] > sq[y]; });
                    int rank = 0;
                    while (rank < s && !(sq[ord[rank]] < e2)) ++rank;
                    DPlrMtx *nd = d2p_node(B, d, q >> 1);
                    DPlrMtx *lr = (q & 1) ? nd->dh.sw : nd->dh.ne;
                    if (lr->dlr.capacity < rank) dplrresizelr(lr, rank);
                    lr->dlr.rank = rank;
                    for (int k = 0; k < rank; ++k) {
                        const int c = ord[k];
                        const double sv = sq[c];
                        lr->dlr.sig[k] = sv;
                        const double *uk = hU.data() + q * ls + (size_t)c * len;
                        const double *mk = hM.data() + q * ls + (size_t)c * len;
                        for (int i = 0; i < len; ++i) {
                            lr->dlr.u[(size_t)k * len + i] = uk[i];
                            lr->dlr.v[(size_t)k * len + i] = mk[i] / sv;
                        }
                    }
                }
                break;
            }
        }
        failed = 0;
    out:
        if (dA) (void)hipFree(dA);
        if (dB) (void)hipFree(dB);
        {
            double *bufs[] = {bf.dY, bf.dZ, bf.dM, bf.dU, bf.dW, bf.dJ, bf.dtau};
            for (double *p : bufs)
                if (p) (void)hipFree(p);
        }
        rocblas_destroy_handle(h);
        if (!failed) *info = 0;
    }
}
