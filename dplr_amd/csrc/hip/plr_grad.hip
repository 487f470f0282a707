// Adjoint of the PLR structure: an outer-product gradient G = alpha * A_ B_^T
// (A_, B_ are N x n blocks of columns) projected onto the factors of the plan,
// i.e. the gradient of <G, A> with respect to the leaves, U, V and sig.  For
// a block U diag(sig) V^T that covers rows of segment q and columns of
// segment q^1 (level-wise layout: U rows on q, V rows on q^1):
//   dU|_q    = alpha * A_|_q (B_|_{q^1}^T V) diag(sig)
//   dV|_{q^1} = alpha * B_|_{q^1} (A_|_q^T U) diag(sig)
//   dsig     = alpha * diag(U^T A_|_q B_|_{q^1}^T V)
//   dD_leaf  = alpha * A_|_leaf B_|_leaf^T
// With y = A^{-1} b and lambda = A^{-T} ybar (a transposed solve), the
// gradient of a loss through y is G = -lambda y^T; through y = A x it is
// G = ybar x^T.  Outputs use the plan's layouts (U for dU and dV, S for dsig,
// the leaves for dD) and are overwritten.  FP64, or FP32 on the plan's FP32
// copies of U, V and sig with FP32 arithmetic.
#include <cstring>

#include "plr_common.h"

template <typename T>
struct GradTab
{
    int L, N;
    int R[PLR_MAXL];
    int col_off[PLR_MAXL + 1]; // first column of level d in the level-wise arrays
    int s_off[PLR_MAXL + 1];   // offset of level d in the S layout
    const T *P[PLR_MAXL]; // reduction operand of the call (U or V)
};

template <typename T>
static void grad_tab(const dplrh_plan *p, const T *U, const T *V, int useU, GradTab<T> *t)
{
    memset(t, 0, sizeof(*t));
    t->L = p->L;
    t->N = p->N;
    for (int d = 0; d < p->L; ++d) {
        t->R[d] = p->R[d];
        t->col_off[d] = (int)(p->uoff[d] / p->N);
        t->s_off[d] = (int)p->soff[d];
        t->P[d] = (useU ? U : V) + p->uoff[d];
    }
    t->col_off[p->L] = (int)(p->uoff[p->L] / p->N);
    t->s_off[p->L] = (int)p->soff[p->L];
}

// T_[c][s*R + k] = P|_s^T X|_s for every segment s of every level (raw sums,
// layout of S per column), P = V (which = 0) or U (which = 1).  One thread
// per (level, segment, rank, column); segments are read in order.
template <typename T>
__global__ void k_grad_reduce(GradTab<T> t, const T *__restrict__ X, int ldx, int ncols,
                              T *__restrict__ out, size_t ostride)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = t.s_off[t.L]; // segments x ranks summed over levels
    if (e >= per * ncols) return;
    const int c = (int)(e / per);
    const int idx = (int)(e % per);
    int d = 0;
    while (idx >= t.s_off[d + 1]) ++d;
    const int loc = idx - t.s_off[d], R = t.R[d];
    const int s = loc / R, k = loc - s * R;
    const int seg = t.N >> (d + 1);
    const T *p = t.P[d] + (size_t)k * t.N + (size_t)s * seg;
    const T *x = X + (size_t)c * ldx + (size_t)s * seg;
    T acc[4] = {0, 0, 0, 0};
    int i = 0;
    for (; i + 4 <= seg; i += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] += p[i + u] * x[i + u];
    }
    for (; i < seg; ++i) acc[0] += p[i] * x[i];
    out[(size_t)c * ostride + idx] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// dU and dV: one thread per (level, row, rank).
template <typename T>
__global__ void k_grad_lr(GradTab<T> t, const T *__restrict__ S, const T *__restrict__ A, int lda,
                          const T *__restrict__ Bm, int ldb, int ncols, T alpha,
                          const T *__restrict__ Ta, const T *__restrict__ Tb, size_t tstride,
                          T *__restrict__ dU, T *__restrict__ dV)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int N = t.N;
    const size_t tot = (size_t)N * t.col_off[t.L];
    if (e >= tot) return;
    const int col = (int)(e / N), i = (int)(e % N); // column of the level-wise array
    int d = 0;
    while (col >= t.col_off[d + 1]) ++d;
    const int k = col - t.col_off[d], R = t.R[d];
    const int seg = N >> (d + 1), q = i / seg;
    const size_t so = t.s_off[d];
    // dU row i (segment q): block with U on q, its V on q^1
    const T sq = S[so + (size_t)q * R + k];
    // dV row i (segment q): block with U on q^1
    const T sp = S[so + (size_t)(q ^ 1) * R + k];
    T gu = 0, gv = 0;
    for (int c = 0; c < ncols; ++c) {
        const T a = A[(size_t)c * lda + i], b = Bm[(size_t)c * ldb + i];
        gu += a * Tb[(size_t)c * tstride + so + (size_t)(q ^ 1) * R + k];
        gv += b * Ta[(size_t)c * tstride + so + (size_t)(q ^ 1) * R + k];
    }
    dU[e] = alpha * sq * gu;
    dV[e] = alpha * sp * gv;
}

// dsig: one thread per (level, segment, rank).
template <typename T>
__global__ void k_grad_sig(GradTab<T> t, int ncols, T alpha, const T *__restrict__ Ta,
                           const T *__restrict__ Tb, size_t tstride, T *__restrict__ dS)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= t.s_off[t.L]) return;
    int d = 0;
    while (idx >= t.s_off[d + 1]) ++d;
    const int loc = idx - t.s_off[d], R = t.R[d];
    const int q = loc / R, k = loc - q * R;
    const size_t so = t.s_off[d];
    T g = 0;
    for (int c = 0; c < ncols; ++c)
        g += Ta[(size_t)c * tstride + idx] * Tb[(size_t)c * tstride + so + (size_t)(q ^ 1) * R + k];
    dS[idx] = alpha * g;
}

// dD: one thread per leaf entry (column-major b x b per leaf).
template <typename T>
__global__ void k_grad_leaf(int N, int b, const T *__restrict__ A, int lda, const T *__restrict__ Bm,
                            int ldb, int ncols, T alpha, T *__restrict__ dD)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * b) return;
    const int l = (int)(e / ((size_t)b * b));
    const int r = (int)(e % ((size_t)b * b));
    const int i = r % b, j = r / b;
    T g = 0;
    for (int c = 0; c < ncols; ++c)
        g += A[(size_t)c * lda + (size_t)l * b + i] * Bm[(size_t)c * ldb + (size_t)l * b + j];
    dD[e] = alpha * g;
}

template <typename T>
static int grad_outer(dplrh_plan *p, const T *U, const T *V, const T *S, int n, T alpha,
                      const T *A, int lda, const T *Bm, int ldb, T *dU, T *dV, T *dS, T *dD,
                      hipStream_t st)
{
    const int N = p->N, L = p->L, b = p->b;
    if (N == 0 || n <= 0) return 0;
    if (lda < N || ldb < N) {
        plr_set_error("leading dimensions must be >= N");
        return -1;
    }
    const size_t nd = (size_t)N * b;
    hipLaunchKernelGGL(k_grad_leaf<T>, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, N, b,
                       A, lda, Bm, ldb, n, alpha, dD);
    PLR_HIP(hipGetLastError());
    if (L == 0) return 0;
    // raw reductions: Ta = U^T A_ per segment, Tb = V^T B_ per segment
    GradTab<T> tv, tu;
    grad_tab(p, U, V, 0, &tv); // P = V
    grad_tab(p, U, V, 1, &tu); // P = U
    const size_t per = p->soff[L];
    T *Tab = nullptr;
    PLR_HIP(hipMallocAsync((void **)&Tab, 2 * per * n * sizeof(T), st));
    T *Ta = Tab, *Tb = Tab + per * n;
    const size_t ne = per * n;
    hipLaunchKernelGGL(k_grad_reduce<T>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, tu,
                       A, lda, n, Ta, per);
    hipLaunchKernelGGL(k_grad_reduce<T>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, tv,
                       Bm, ldb, n, Tb, per);
    const size_t nl = (size_t)N * tv.col_off[L];
    hipLaunchKernelGGL(k_grad_lr<T>, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st, tv, S,
                       A, lda, Bm, ldb, n, alpha, Ta, Tb, per, dU, dV);
    hipLaunchKernelGGL(k_grad_sig<T>, dim3((unsigned)((per + 255) / 256)), dim3(256), 0, st, tv, n,
                       alpha, Ta, Tb, per, dS);
    PLR_HIP(hipGetLastError());
    PLR_HIP(hipFreeAsync(Tab, st));
    return 0;
}

extern "C" int dplrh_grad_outer(dplrh_plan *p, int n, double alpha, const double *A, int lda,
                                const double *Bm, int ldb, double *dU, double *dV, double *dS,
                                double *dD, void *stream)
{
    return grad_outer<double>(p, p->U, p->V, p->S, n, alpha, A, lda, Bm, ldb, dU, dV, dS, dD,
                              (hipStream_t)stream);
}

// FP32: the factors' FP32 copies (dplrh_sp_factor type 1), FP32 arithmetic.
extern "C" int dplrh_grad_outer_f32(dplrh_plan *p, int n, float alpha, const float *A, int lda,
                                    const float *Bm, int ldb, float *dU, float *dV, float *dS,
                                    float *dD, void *stream)
{
    const plr_sp *s = p->sp[0];
    if (!s) {
        plr_set_error("FP32 gradient before the plan's FP32 factors (type 1) exist");
        return -4;
    }
    return grad_outer<float>(p, (const float *)s->U, (const float *)s->V, s->S, n, alpha, A, lda,
                             Bm, ldb, dU, dV, dS, dD, (hipStream_t)stream);
}
