// Explicit PLR inverse in the reference's structure (dplrgeinv, reference
// src/dplrinv.c:93) and the rank-k update (dplrgeprk, src/dplrprk.c:21),
// computed on the GPU from the factored form of plr_solve.hip.
//
// For a node at depth d with K = I + Z^T inv(D) W S (see plr_solve.hip) the
// reference's X equals K inv(S), so its Y = inv(X) = S inv(K) and
//   B.ne = -Ut|_top (S_top.*Kinv_11) Vt|_bot^T   B.sw = -Ut|_bot (S_bot.*Kinv_22) Vt|_top^T
//   correction into the upper subtree: u = -Ut|_top (S_top.*Kinv_12), v = Vt|_top
//   correction into the lower subtree: u = -Ut|_bot (S_bot.*Kinv_21), v = Vt|_bot
// where Ut_d = inv(D_{>d}) U_d (the factorization's Ut) and
// Vt_d = inv(D_{>d})^T V_d (computed here by a transposed top-down sweep).
// The reference then normalises columns per block, zeroes sig below the
// per-level share of eps (leaving such columns unnormalised), and appends the
// corrections to every low-rank block of the subtree (dplrgeprk), adding the
// dense part to the leaves.  Every one of those values is produced here; the
// host only reshapes B and copies.
#include <cmath>
#include <cstring>
#include <vector>

#include "plr_common.h"

// Sum of squares of every column over every leaf segment:
// out[c*nleaf + l] = sum_{i < b} X[c*ldx + l*b + i]^2
__global__ void k_leafsq(const double *__restrict__ X, int ldx, int ncols, int nleaf, int b,
                         double *__restrict__ out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)ncols * nleaf) return;
    const int c = (int)(e / nleaf), l = (int)(e % nleaf);
    const double *x = X + (size_t)c * ldx + (size_t)l * b;
    double s = 0.0;
    for (int i = 0; i < b; ++i) s += x[i] * x[i];
    out[e] = s;
}

// Norm of column c over segment q of level d (pairwise sum of leaf sums).
__device__ double seg_norm(const double *sq, int nleaf, int L, int c, int d, int q)
{
    const int per = 1 << (L - 1 - d); // leaves per segment
    const double *p = sq + (size_t)c * nleaf + (size_t)q * per;
    double s = 0.0;
    for (int j = 0; j < per; ++j) s += p[j];
    return sqrt(s);
}

// Reference rule (src/dplrinv.c:333-348, src/dplrprk.c:113-136): scale
// sc = f * f1 * f2; below thr the value is zeroed and the columns stay as
// they are, otherwise the columns are divided by f1 and f2 (u also takes the
// sign of alpha).
__device__ __forceinline__ void norm_rule(double f, double f1, double f2, double thr, double sgn,
                                          double *sig, double *g1, double *g2)
{
    const double sc = f * f1 * f2;
    if (sc < thr) {
        *sig = 0.0;
        *g1 = 1.0;
        *g2 = 1.0;
    } else {
        *sig = sc;
        *g1 = sgn / f1;
        *g2 = 1.0 / f2;
    }
}

struct InvLevels
{
    int L, N, b, nleaf;
    int R[PLR_MAXL];
    int coff[PLR_MAXL][PLR_MAXL]; // column of the correction from level d in level d' buffers
    int Rt[PLR_MAXL];             // columns of the level d' buffers
    int col[PLR_MAXL];            // first column of level d in the level-wise arrays
    double eps2[PLR_MAXL];
};

// Per (segment q of level d', column c of the level d' buffer): scales of the
// u and v parts and sig.  Own columns come from OU/VT of level d'; the
// correction columns from level d < d' use CU_d/VT_d, normalised first over
// the level-d segment (giving the d2 scale that dplrgeprk receives) and then
// per block.  A column of the true rank whose scale is not finite (a zero
// norm with a zero threshold: the reference divides 0 by 0 there) flags the
// node whose Woodbury step made it in nanorg[(1 << depth) - 1 + node].
__global__ void k_inv_scales(InvLevels a, int dp, const double *sq_ou, const double *sq_cu,
                             const double *sq_vt, const int *rank, double *gu, double *gv,
                             double *sig, int *nanorg)
{
    const int nseg = 2 << dp, Rt = a.Rt[dp];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nseg * Rt) return;
    const int q = e / Rt, c = e % Rt;
    double s_, g1, g2;
    if (c < a.R[dp]) {
        const int col = a.col[dp] + c;
        const double f1 = seg_norm(sq_ou, a.nleaf, a.L, col, dp, q);
        const double f2 = seg_norm(sq_vt, a.nleaf, a.L, col, dp, q ^ 1);
        norm_rule(1.0, f1, f2, a.eps2[dp], 1.0, &s_, &g1, &g2);
        if (c < rank[(2 << dp) - 2 + q] && !(isfinite(g1) && isfinite(g2)))
            nanorg[(1 << dp) - 1 + (q >> 1)] = 1;
    } else {
        int d = dp - 1;
        while (d >= 0 && !(c >= a.coff[dp][d] && c < a.coff[dp][d] + a.R[d])) --d;
        if (d < 0) { // unreachable for a consistent table
            gu[e] = gv[e] = sig[e] = 0.0;
            return;
        }
        const int cp = c - a.coff[dp][d], col = a.col[d] + cp;
        const int s = q >> (dp - d);
        const double F1 = seg_norm(sq_cu, a.nleaf, a.L, col, d, s);
        const double F2 = seg_norm(sq_vt, a.nleaf, a.L, col, d, s);
        double d2, G1, G2;
        norm_rule(1.0, F1, F2, a.eps2[d], 1.0, &d2, &G1, &G2);
        const double f1 = seg_norm(sq_cu, a.nleaf, a.L, col, dp, q) * G1;
        const double f2 = seg_norm(sq_vt, a.nleaf, a.L, col, dp, q ^ 1) * G2;
        norm_rule(d2, f1, f2, a.eps2[dp], 1.0, &s_, &g1, &g2);
        g1 *= G1;
        g2 *= G2;
        if (cp < rank[(2 << d) - 2 + (s ^ 1)] && !(isfinite(g1) && isfinite(g2)))
            nanorg[(1 << d) - 1 + (s >> 1)] = 1;
    }
    gu[e] = g1;
    gv[e] = g2;
    sig[e] = s_;
}

// Leaf factor of the correction columns of level d: G1 * d2 * G2 per
// (segment s of level d, column); zero for the padding columns beyond the
// true rank k of the correction (the rank of the block on the sibling
// segment).
__global__ void k_inv_leaf_factor(InvLevels a, int d, const double *sq_cu, const double *sq_vt,
                                  const int *rank, double *lf, int *nanorg)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nseg = 2 << d, R = a.R[d];
    if (e >= nseg * R) return;
    const int s = e / R, cp = e % R, col = a.col[d] + cp;
    if (cp >= rank[(2 << d) - 2 + (s ^ 1)]) {
        lf[e] = 0.0;
        return;
    }
    const double F1 = seg_norm(sq_cu, a.nleaf, a.L, col, d, s);
    const double F2 = seg_norm(sq_vt, a.nleaf, a.L, col, d, s);
    double d2, G1, G2;
    norm_rule(1.0, F1, F2, a.eps2[d], 1.0, &d2, &G1, &G2);
    if (!(isfinite(G1) && isfinite(G2))) nanorg[(1 << d) - 1 + (s >> 1)] = 1;
    lf[e] = G1 * d2 * G2;
}

// Level d' output: BU[(q*Rt + c)*len + i] = srcU_c[q*len + i] * gu,
// BV[...] = srcV_c[(q^1)*len + i] * gv (the V part of the block whose U is
// on segment q lives on the sibling segment).
__global__ void k_inv_gather(int N, int nseg, int Rt, const double *const *srcU,
                             const double *const *srcV, const double *gu, const double *gv,
                             double *BU, double *BV)
{
    const int len = N / nseg;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * Rt) return;
    const int i = (int)(e % len);
    const size_t qc = e / len;
    const int c = (int)(qc % Rt), q = (int)(qc / Rt);
    const double *su = srcU[c], *sv = srcV[c];
    BU[e] = su[(size_t)q * len + i] * gu[qc];
    BV[e] = sv[(size_t)(q ^ 1) * len + i] * gv[qc];
}

// Own and correction coefficient blocks of level d from the capacitance
// inverses: W_own[c][s][k] = -S_s[k] Kinv[sb R + k][sb R + c],
// W_cor[c][s][k] = -S_s[k] Kinv[sb R + k][(1-sb) R + c], sb = s & 1.
__global__ void k_inv_wblocks(const double *Kinv, const double *S, int R, int nseg, double *Wown,
                              double *Wcor)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = nseg * R; // one column
    if (e >= per * R) return;
    const int c = e / per, r = e % per, s = r / R, k = r % R;
    const int n2 = 2 * R, sb = s & 1, node = s >> 1;
    const double *K = Kinv + (size_t)node * n2 * n2;
    const double sk = -S[(size_t)s * R + k];
    Wown[e] = sk * K[(size_t)(sb * R + c) * n2 + sb * R + k];
    Wcor[e] = sk * K[(size_t)((1 - sb) * R + c) * n2 + sb * R + k];
}

// B leaves: Dinv + sum over levels d of CU_d diag(lf_d) VT_d^T on the leaf.
__global__ void k_inv_leaves(InvLevels a, const double *Dinv, const double *CU, const double *VT,
                             const double *lf, const size_t *lfoff, double *BD)
{
    const int b = a.b, N = a.N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * b) return;
    const int l = (int)(e / ((size_t)b * b));
    const int r = (int)(e % ((size_t)b * b));
    const int i = r % b, j = r / b; // column-major (i, j)
    double v = Dinv[e];
    const int row_i = l * b + i, row_j = l * b + j;
    for (int d = 0; d < a.L; ++d) {
        const int s = l >> (a.L - 1 - d);
        const double *f = lf + lfoff[d] + (size_t)s * a.R[d];
        for (int cp = 0; cp < a.R[d]; ++cp) {
            const size_t col = (size_t)(a.col[d] + cp) * N;
            v += CU[col + row_i] * f[cp] * VT[col + row_j];
        }
    }
    BD[e] = v;
}

// ---------------------------------------------------------------------------

static double eps2_of(int d, int L, double eps)
{
    double e = eps;
    for (int k = 0; k < d; ++k) e /= 4.0;
    return (d == L - 1) ? e / 2.0 : e / 4.0;
}

// Compute the explicit inverse of a factored plan and copy it to host
// buffers: for every level d', hBU[d'] and hBV[d'] hold N*Rt[d'] doubles
// laid out [segment q][column c][row], hBsig[d'] holds nseg*Rt[d'] doubles;
// hBD holds the N/b leaves (column-major); hnan (2^L ints, may be null) the
// flags of the nodes whose Woodbury step divided 0 by 0.  Rt and coff as in
// InvLevels.
extern "C" int dplrh_geinv(dplrh_plan *p, double eps, const int *Rt, int coff[][DPLRH_MAXL],
                           double **hBU, double **hBV, double **hBsig, double *hBD, int *hnan,
                           void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (!p->f.done) {
        plr_set_error("explicit inverse before a successful factorization");
        return -4;
    }
    const int N = p->N, L = p->L, b = p->b, nleaf = N / b;
    const size_t ctot = p->uoff[L] / (N ? N : 1);
    const size_t nd = (size_t)N * b;
    InvLevels a;
    memset(&a, 0, sizeof(a));
    a.L = L;
    a.N = N;
    a.b = b;
    a.nleaf = nleaf;
    for (int d = 0; d < L; ++d) {
        a.R[d] = p->R[d];
        a.Rt[d] = Rt[d];
        a.col[d] = (int)(p->uoff[d] / N);
        a.eps2[d] = eps2_of(d, L, eps);
        for (int e = 0; e < L; ++e) a.coff[d][e] = coff[d][e];
    }
    if (hnan) memset(hnan, 0, ((size_t)1 << L) * sizeof(int));
    if (L == 0) {
        PLR_HIP(hipMemcpy(hBD, p->f.Dinv, nd * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    }
    double *VT = nullptr, *Y = nullptr, *OU = nullptr, *CU = nullptr, *sq = nullptr;
    double *gu = nullptr, *gv = nullptr, *sg = nullptr, *lf = nullptr, *BU = nullptr, *BV = nullptr;
    double *Wown = nullptr, *Wcor = nullptr, *BD = nullptr;
    size_t *lfoff = nullptr;
    int *nanorg = nullptr;
    const double **tab = nullptr;
    int rc = 0;
    int maxRt = 1, maxR = 1;
    for (int d = 0; d < L; ++d) {
        if (Rt[d] > maxRt) maxRt = Rt[d];
        if (p->R[d] > maxR) maxR = p->R[d];
    }
    size_t lftot = 0;
    std::vector<size_t> hlfoff(L);
    for (int d = 0; d < L; ++d) {
        hlfoff[d] = lftot;
        lftot += ((size_t)2 << d) * p->R[d];
    }
    auto A = [&](void **ptr, size_t bytes) -> int {
        PLR_HIP(hipMalloc(ptr, bytes ? bytes : 8));
        return 0;
    };
    if ((rc = A((void **)&VT, ctot * N * 8)) || (rc = A((void **)&Y, ctot * N * 8)) ||
        (rc = A((void **)&OU, ctot * N * 8)) || (rc = A((void **)&CU, ctot * N * 8)) ||
        (rc = A((void **)&sq, 3 * ctot * nleaf * 8)) ||
        (rc = A((void **)&gu, ((size_t)2 << (L - 1)) * maxRt * 8)) ||
        (rc = A((void **)&gv, ((size_t)2 << (L - 1)) * maxRt * 8)) ||
        (rc = A((void **)&sg, ((size_t)2 << (L - 1)) * maxRt * 8)) || (rc = A((void **)&lf, lftot * 8)) ||
        (rc = A((void **)&lfoff, L * sizeof(size_t))) || (rc = A((void **)&BU, (size_t)N * maxRt * 8)) ||
        (rc = A((void **)&BV, (size_t)N * maxRt * 8)) ||
        (rc = A((void **)&Wown, ((size_t)2 << (L - 1)) * maxR * maxR * 8)) ||
        (rc = A((void **)&Wcor, ((size_t)2 << (L - 1)) * maxR * maxR * 8)) ||
        (rc = A((void **)&BD, nd * 8)) || (rc = A((void **)&tab, 2 * (size_t)maxRt * sizeof(double *))) ||
        (rc = A((void **)&nanorg, ((size_t)1 << L) * sizeof(int))))
        goto done;
    PLR_HIP(hipMemsetAsync(nanorg, 0, ((size_t)1 << L) * sizeof(int), st));
    PLR_HIP(hipMemcpyAsync(lfoff, hlfoff.data(), L * sizeof(size_t), hipMemcpyHostToDevice, st));

    // 1. Vt: V with F_{d'}^T applied for every d' above, then Dinv^T.
    PLR_HIP(hipMemcpyAsync(Y, p->V, ctot * N * 8, hipMemcpyDeviceToDevice, st));
    for (int dp = 1; dp < L && !rc; ++dp) {
        const int off = a.col[dp];
        if (p->R[dp] == 0 || off == 0) continue;
        if ((rc = plr_solve_reserve(p, off))) break;
        const double *Ud = p->f.Ut + p->uoff[dp], *Vd = p->V + p->uoff[dp];
        if ((rc = plr_lvl_reduce(p, dp, Ud, Y, N, off, p->sw.T, p->sw.t_per_col, st))) break;
        if ((rc = plr_lvl_cap(p, dp, p->sw.T, p->sw.t_per_col, p->sw.W, off, 1, st))) break;
        rc = plr_lvl_expand(p, dp, Vd, p->sw.W, p->sw.t_per_col, nullptr, 0, nullptr, N, 1.0, 1.0,
                            Y, N, off, st);
    }
    if (rc) goto done;
    if ((rc = plr_lvl_expand(p, -1, nullptr, nullptr, 0, p->f.Dinv, 1, Y, N, 1.0, 0.0, VT, N,
                             (int)ctot, st)))
        goto done;

    // 2. own (OU) and correction (CU) u-factors of every level
    for (int d = 0; d < L && !rc; ++d) {
        const int R = p->R[d], nseg = 2 << d;
        if (R == 0) continue;
        const int ne = nseg * R * R;
        hipLaunchKernelGGL(k_inv_wblocks, dim3((ne + 255) / 256), dim3(256), 0, st,
                           p->f.Kinv + p->f.koff[d], p->S + p->soff[d], R, nseg, Wown, Wcor);
        PLR_HIP(hipGetLastError());
        const double *Ud = p->f.Ut + p->uoff[d];
        if ((rc = plr_lvl_expand(p, d, Ud, Wown, (size_t)nseg * R, nullptr, 0, nullptr, N, 1.0, 0.0,
                                 OU + p->uoff[d], N, R, st)))
            break;
        rc = plr_lvl_expand(p, d, Ud, Wcor, (size_t)nseg * R, nullptr, 0, nullptr, N, 1.0, 0.0,
                            CU + p->uoff[d], N, R, st);
    }
    if (rc) goto done;

    // 3. leaf-segment sums of squares of OU, CU, VT
    {
        const size_t ne = ctot * nleaf;
        if (ne) {
            hipLaunchKernelGGL(k_leafsq, dim3((ne + 255) / 256), dim3(256), 0, st, OU, N, (int)ctot,
                               nleaf, b, sq);
            hipLaunchKernelGGL(k_leafsq, dim3((ne + 255) / 256), dim3(256), 0, st, CU, N, (int)ctot,
                               nleaf, b, sq + ne);
            hipLaunchKernelGGL(k_leafsq, dim3((ne + 255) / 256), dim3(256), 0, st, VT, N, (int)ctot,
                               nleaf, b, sq + 2 * ne);
            PLR_HIP(hipGetLastError());
        }
        const double *sq_ou = sq, *sq_cu = sq + ne, *sq_vt = sq + 2 * ne;
        for (int d = 0; d < L; ++d) {
            const int nl = (2 << d) * p->R[d];
            if (nl == 0) continue;
            hipLaunchKernelGGL(k_inv_leaf_factor, dim3((nl + 255) / 256), dim3(256), 0, st, a, d,
                               sq_cu, sq_vt, p->rank, lf + hlfoff[d], nanorg);
            PLR_HIP(hipGetLastError());
        }
        // 4. per level: scales, gathered factors, download
        std::vector<const double *> htab(2 * (size_t)maxRt);
        for (int dp = 0; dp < L && !rc; ++dp) {
            const int nseg = 2 << dp, R = Rt[dp];
            if (R == 0) continue;
            const int ns = nseg * R;
            hipLaunchKernelGGL(k_inv_scales, dim3((ns + 255) / 256), dim3(256), 0, st, a, dp, sq_ou,
                               sq_cu, sq_vt, p->rank, gu, gv, sg, nanorg);
            PLR_HIP(hipGetLastError());
            for (int c = 0; c < R; ++c) {
                const double *su = nullptr, *sv = nullptr;
                if (c < p->R[dp]) {
                    su = OU + p->uoff[dp] + (size_t)c * N;
                    sv = VT + p->uoff[dp] + (size_t)c * N;
                } else {
                    for (int d = dp - 1; d >= 0; --d)
                        if (c >= coff[dp][d] && c < coff[dp][d] + p->R[d]) {
                            su = CU + p->uoff[d] + (size_t)(c - coff[dp][d]) * N;
                            sv = VT + p->uoff[d] + (size_t)(c - coff[dp][d]) * N;
                        }
                }
                htab[c] = su;
                htab[R + c] = sv;
            }
            PLR_HIP(hipMemcpyAsync(tab, htab.data(), 2 * (size_t)R * sizeof(double *),
                                   hipMemcpyHostToDevice, st));
            const size_t ne2 = (size_t)N * R;
            hipLaunchKernelGGL(k_inv_gather, dim3((ne2 + 255) / 256), dim3(256), 0, st, N, nseg, R,
                               tab, tab + R, gu, gv, BU, BV);
            PLR_HIP(hipGetLastError());
            PLR_HIP(hipMemcpyAsync(hBU[dp], BU, ne2 * 8, hipMemcpyDeviceToHost, st));
            PLR_HIP(hipMemcpyAsync(hBV[dp], BV, ne2 * 8, hipMemcpyDeviceToHost, st));
            PLR_HIP(hipMemcpyAsync(hBsig[dp], sg, (size_t)ns * 8, hipMemcpyDeviceToHost, st));
            PLR_HIP(hipStreamSynchronize(st));
        }
        if (rc) goto done;
        // 5. leaves
        hipLaunchKernelGGL(k_inv_leaves, dim3((nd + 255) / 256), dim3(256), 0, st, a, p->f.Dinv, CU,
                           VT, lf, lfoff, BD);
        PLR_HIP(hipGetLastError());
        PLR_HIP(hipMemcpyAsync(hBD, BD, nd * 8, hipMemcpyDeviceToHost, st));
        if (hnan)
            PLR_HIP(hipMemcpyAsync(hnan, nanorg, ((size_t)1 << L) * sizeof(int),
                                   hipMemcpyDeviceToHost, st));
        PLR_HIP(hipStreamSynchronize(st));
    }
done:
    void *bufs[] = {VT,   Y,    OU,   CU,  sq, gu,          gv,    sg,
                    lf,   lfoff, BU,  BV,  Wown, Wcor,      BD,    (void *)tab, nanorg};
    for (void *x : bufs)
        if (x) (void)hipFree(x);
    return rc;
}

// ---------------------------------------------------------------------------
// rank-k update A += alpha * u diag(s) v^T (reference src/dplrprk.c:21)

__global__ void k_prk_scales(int L, int nleaf, int k, int d, double alpha, const double *s,
                             const double *sq_u, const double *sq_v, double eps2, double *gu,
                             double *gv, double *sig)
{
    const int nseg = 2 << d;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nseg * k) return;
    const int q = e / k, c = e % k;
    const double f1 = seg_norm(sq_u, nleaf, L, c, d, q);
    const double f2 = seg_norm(sq_v, nleaf, L, c, d, q ^ 1);
    double sg, g1, g2;
    norm_rule(fabs(alpha) * s[c], f1, f2, eps2, alpha > 0.0 ? 1.0 : -1.0, &sg, &g1, &g2);
    gu[e] = g1;
    gv[e] = g2;
    sig[e] = sg;
}

// D_l += alpha * sum_c u_c|_l s_c v_c|_l^T (dense leaf update, BLAS dgemm 'N','T').
__global__ void k_prk_leaves(int N, int b, int k, double alpha, const double *u, const double *v,
                             const double *s, double *D)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * b) return;
    const int l = (int)(e / ((size_t)b * b));
    const int r = (int)(e % ((size_t)b * b));
    const int i = r % b, j = r / b;
    double acc = 0.0;
    for (int c = 0; c < k; ++c) acc += u[(size_t)c * N + l * b + i] * s[c] * v[(size_t)c * N + l * b + j];
    D[e] += alpha * acc;
}

// hu, hv: N x k column-major (ld N); hs: k scales; hD: leaves (in/out).
// Per level d: hUn[d], hVn[d] receive N*k doubles laid out [q][c][row] and
// hsig[d] nseg*k doubles, ready to be appended to each block.
// Device buffers of dplrh_geprk, kept from call to call while the calling
// thread's pool is on (the rank-k corrections of dplrgeppmm: a thousand
// calls at N=65536, each of which would otherwise allocate and free its
// eleven buffers); grown when a call needs more.
namespace {
struct PrkPool
{
    bool on;
    int dev;
    void *p[11];
    size_t cap[11];
};
thread_local PrkPool g_prk;

void prk_pool_free()
{
    for (int i = 0; i < 11; ++i) {
        if (g_prk.p[i]) (void)hipFree(g_prk.p[i]);
        g_prk.p[i] = nullptr;
        g_prk.cap[i] = 0;
    }
}
} // namespace

extern "C" void dplrh_prk_pool(int on)
{
    if (!on) prk_pool_free();
    g_prk.on = on != 0;
    if (on) (void)hipGetDevice(&g_prk.dev);
}

extern "C" int dplrh_geprk(int N, int L, int b, int k, double alpha, double eps, const double *hu,
                           const double *hv, const double *hs, double *hD, double **hUn,
                           double **hVn, double **hsig, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int nleaf = N / b;
    const size_t nk = (size_t)N * k, nd = (size_t)N * b;
    double *u = nullptr, *v = nullptr, *s = nullptr, *D = nullptr, *sq = nullptr, *gu = nullptr;
    double *gv = nullptr, *sg = nullptr, *BU = nullptr, *BV = nullptr;
    const double **tab = nullptr;
    int rc = 0;
    const size_t maxseg = L > 0 ? ((size_t)2 << (L - 1)) : 1;
    const size_t Lc = L > 0 ? (size_t)L : 1; // every level's factors and sig stay until one wait
    bool pooled = g_prk.on;
    if (pooled) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (dev != g_prk.dev) pooled = false;
    }
    int slot = 0;
    auto A = [&](void **ptr, size_t bytes) -> int {
        bytes = bytes ? bytes : 8;
        if (pooled) {
            const int i = slot++;
            if (g_prk.cap[i] < bytes) {
                if (g_prk.p[i]) (void)hipFree(g_prk.p[i]);
                g_prk.p[i] = nullptr;
                g_prk.cap[i] = 0;
                PLR_HIP(hipMalloc(&g_prk.p[i], bytes));
                g_prk.cap[i] = bytes;
            }
            *ptr = g_prk.p[i];
            return 0;
        }
        PLR_HIP(hipMalloc(ptr, bytes));
        return 0;
    };
    if ((rc = A((void **)&u, nk * 8)) || (rc = A((void **)&v, nk * 8)) || (rc = A((void **)&s, k * 8)) ||
        (rc = A((void **)&D, nd * 8)) || (rc = A((void **)&sq, 2 * (size_t)k * nleaf * 8)) ||
        (rc = A((void **)&gu, maxseg * k * 8)) || (rc = A((void **)&gv, maxseg * k * 8)) ||
        (rc = A((void **)&sg, Lc * maxseg * k * 8)) || (rc = A((void **)&BU, Lc * nk * 8)) ||
        (rc = A((void **)&BV, Lc * nk * 8)) || (rc = A((void **)&tab, 2 * (size_t)k * sizeof(double *))))
        goto done;
    PLR_HIP(hipMemcpyAsync(u, hu, nk * 8, hipMemcpyHostToDevice, st));
    PLR_HIP(hipMemcpyAsync(v, hv, nk * 8, hipMemcpyHostToDevice, st));
    PLR_HIP(hipMemcpyAsync(s, hs, (size_t)k * 8, hipMemcpyHostToDevice, st));
    PLR_HIP(hipMemcpyAsync(D, hD, nd * 8, hipMemcpyHostToDevice, st));
    {
        const size_t ne = (size_t)k * nleaf;
        hipLaunchKernelGGL(k_leafsq, dim3((ne + 255) / 256), dim3(256), 0, st, u, N, k, nleaf, b, sq);
        hipLaunchKernelGGL(k_leafsq, dim3((ne + 255) / 256), dim3(256), 0, st, v, N, k, nleaf, b,
                           sq + ne);
        PLR_HIP(hipGetLastError());
        std::vector<const double *> htab(2 * (size_t)k);
        for (int c = 0; c < k; ++c) {
            htab[c] = u + (size_t)c * N;
            htab[k + c] = v + (size_t)c * N;
        }
        PLR_HIP(hipMemcpyAsync(tab, htab.data(), 2 * (size_t)k * sizeof(double *),
                               hipMemcpyHostToDevice, st));
        // each level into its own buffers, one wait for all of them
        for (int d = 0; d < L; ++d) {
            const int nseg = 2 << d, ns = nseg * k;
            double *sgd = sg + (size_t)d * maxseg * k, *BUd = BU + (size_t)d * nk, *BVd = BV + (size_t)d * nk;
            hipLaunchKernelGGL(k_prk_scales, dim3((ns + 255) / 256), dim3(256), 0, st, L, nleaf, k,
                               d, alpha, s, sq, sq + ne, eps2_of(d, L, eps), gu, gv, sgd);
            hipLaunchKernelGGL(k_inv_gather, dim3((nk + 255) / 256), dim3(256), 0, st, N, nseg, k,
                               tab, tab + k, gu, gv, BUd, BVd);
            PLR_HIP(hipGetLastError());
            PLR_HIP(hipMemcpyAsync(hUn[d], BUd, nk * 8, hipMemcpyDeviceToHost, st));
            PLR_HIP(hipMemcpyAsync(hVn[d], BVd, nk * 8, hipMemcpyDeviceToHost, st));
            PLR_HIP(hipMemcpyAsync(hsig[d], sgd, (size_t)ns * 8, hipMemcpyDeviceToHost, st));
        }
        hipLaunchKernelGGL(k_prk_leaves, dim3((nd + 255) / 256), dim3(256), 0, st, N, b, k, alpha,
                           u, v, s, D);
        PLR_HIP(hipGetLastError());
        PLR_HIP(hipMemcpyAsync(hD, D, nd * 8, hipMemcpyDeviceToHost, st));
        PLR_HIP(hipStreamSynchronize(st));
    }
done:
    if (pooled) {
        // the next call on this thread reuses them: nothing may be in flight
        if (rc) (void)hipStreamSynchronize(st);
        return rc;
    }
    void *bufs[] = {u, v, s, D, sq, gu, gv, sg, BU, BV, (void *)tab};
    for (void *x : bufs)
        if (x) (void)hipFree(x);
    return rc;
}
