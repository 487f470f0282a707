// Compression of a dense host matrix into a PLR tree: the reference's
// dplrd2gep (src/dplrd2p.c:187) with the same tree shaping, the same split
// of eps over the levels (eps/2 at the last level, eps/4 above, per level of
// recursion) and the same rank rule (the number of leading singular values
// >= the block's share of eps, src/dplrd2p.c:126-127).  The singular value
// decompositions run on the GPU: every off-diagonal block of one level is
// gathered into a strided batch and factored by one call of rocSOLVER's
// batched SVD; U and V^T of the kept singular triplets come back to the
// tree.  Singular vectors may differ in sign from LAPACK's (u_k v_k^T does
// not).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

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

#define D2P_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            fprintf(stderr, "dplrd2gep: %s: %s\n", #call, hipGetErrorString(e_));                  \
            goto out;                                                                              \
        }                                                                                          \
    } while (0)

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
        double *dA = nullptr, *dB = nullptr, *dS = nullptr, *dU = nullptr, *dV = nullptr,
               *dE = nullptr;
        rocblas_int *dinfo = nullptr;
        rocblas_handle h = nullptr;
        int failed = 1;
        const size_t half = (size_t)m * m / 2; // every level's blocks cover half of A
        std::vector<double> hS, hU, hVt;
        std::vector<rocblas_int> hinfo;
        if (rocblas_create_handle(&h) != rocblas_status_success) {
            fprintf(stderr, "%s: rocblas_create_handle failed\n", __func__);
            return;
        }
        D2P_HIP(hipMalloc(&dA, (size_t)m * m * sizeof(double)));
        D2P_HIP(hipMalloc(&dB, half * sizeof(double)));
        D2P_HIP(hipMalloc(&dU, half * sizeof(double)));
        D2P_HIP(hipMalloc(&dV, half * sizeof(double)));
        D2P_HIP(hipMalloc(&dS, (size_t)m * sizeof(double)));
        D2P_HIP(hipMalloc(&dE, (size_t)m * sizeof(double)));
        D2P_HIP(hipMalloc(&dinfo, (size_t)m * sizeof(rocblas_int)));
        D2P_HIP(hipMemcpy2D(dA, (size_t)m * sizeof(double), A, (size_t)lda * sizeof(double),
                            (size_t)m * sizeof(double), m, hipMemcpyHostToDevice));
        for (int d = 0; d < nlvl; ++d) {
            const int len = m >> (d + 1), nblk = 2 << d;
            const size_t per = (size_t)len * len, tot = per * nblk;
            const double e2 = d2p_eps(eps, d, nlvl);
            hipLaunchKernelGGL(k_d2p_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0,
                               dA, (size_t)m, len, nblk, dB);
            D2P_HIP(hipGetLastError());
            rocblas_status rs = rocsolver_dgesvd_strided_batched(
                h, rocblas_svect_singular, rocblas_svect_singular, len, len, dB, len, per, dS, len,
                dU, len, per, dV, len, per, dE, len, rocblas_outofplace, dinfo, nblk);
            if (rs != rocblas_status_success) {
                fprintf(stderr, "%s: batched SVD failed (%d)\n", __func__, (int)rs);
                goto out;
            }
            hS.resize((size_t)nblk * len);
            hinfo.resize(nblk);
            D2P_HIP(hipMemcpy(hS.data(), dS, hS.size() * sizeof(double), hipMemcpyDeviceToHost));
            D2P_HIP(hipMemcpy(hinfo.data(), dinfo, nblk * sizeof(rocblas_int),
                              hipMemcpyDeviceToHost));
            for (int q = 0; q < nblk; ++q) {
                if (hinfo[q] != 0) {
                    fprintf(stderr, "%s: SVD did not converge (%d)\n", __func__, hinfo[q]);
                    *info = hinfo[q];
                    goto out;
                }
                const double *s = hS.data() + (size_t)q * len;
                int rank = 0;
                while (rank < len && !(s[rank] < e2)) ++rank;
                DPlrMtx *nd = d2p_node(B, d, q >> 1);
                DPlrMtx *lr = (q & 1) ? nd->dh.sw : nd->dh.ne;
                if (lr->dlr.capacity < rank) dplrresizelr(lr, rank);
                lr->dlr.rank = rank;
                if (rank == 0) continue;
                memcpy(lr->dlr.sig, s, (size_t)rank * sizeof(double));
                D2P_HIP(hipMemcpy(lr->dlr.u, dU + (size_t)q * per,
                                  (size_t)rank * len * sizeof(double), hipMemcpyDeviceToHost));
                // the first `rank` rows of V^T (column-major len x len), transposed
                hVt.resize((size_t)rank * len);
                D2P_HIP(hipMemcpy2D(hVt.data(), (size_t)rank * sizeof(double),
                                    dV + (size_t)q * per, (size_t)len * sizeof(double),
                                    (size_t)rank * sizeof(double), len, hipMemcpyDeviceToHost));
                for (int k = 0; k < rank; ++k)
                    for (int i = 0; i < len; ++i)
                        lr->dlr.v[(size_t)k * len + i] = hVt[(size_t)i * rank + k];
            }
        }
        failed = 0;
    out:
        if (dA) (void)hipFree(dA);
        if (dB) (void)hipFree(dB);
        if (dU) (void)hipFree(dU);
        if (dV) (void)hipFree(dV);
        if (dS) (void)hipFree(dS);
        if (dE) (void)hipFree(dE);
        if (dinfo) (void)hipFree(dinfo);
        rocblas_destroy_handle(h);
        if (!failed) *info = 0;
    }
}
