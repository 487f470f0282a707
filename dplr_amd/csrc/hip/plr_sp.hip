// Single-precision copies of a plan's factors: U, V, sig, the leaves and,
// once the FP64 factorization exists, Ut, the leaves' inverses and the
// capacitance inverses, in FP32 (type 1) or with the factor matrices in BF16
// and sig and the capacitance inverses in FP32 (type 2).  The pivoted leaf
// LU and every capacitance LU run in FP64 (dplrh_factor); only the storage
// and the arithmetic of the products and solves are single precision.
// Layouts are those of the FP64 arrays (level-wise columns at uoff, sig at
// soff, leaves column-major, inv(K) column-major at koff).
#include <cstdint>
#include <cstring>

#include "plr_common.h"

// round to nearest even; NaN stays NaN
__device__ __forceinline__ uint16_t f2bf(float x)
{
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ void k_sp_f32(const double *__restrict__ src, size_t n, float *__restrict__ dst)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

__global__ void k_sp_bf16(const double *__restrict__ src, size_t n, uint16_t *__restrict__ dst)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = f2bf((float)src[i]);
}

static int sp_convert(const double *src, size_t n, int bf16, void *dst, hipStream_t st)
{
    if (n == 0) return 0;
    const unsigned blocks = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    if (bf16)
        hipLaunchKernelGGL(k_sp_bf16, dim3(blocks), dim3(256), 0, st, src, n, (uint16_t *)dst);
    else
        hipLaunchKernelGGL(k_sp_f32, dim3(blocks), dim3(256), 0, st, src, n, (float *)dst);
    PLR_HIP(hipGetLastError());
    return 0;
}

static void sp_free(plr_sp *s)
{
    if (!s) return;
    void *bufs[] = {s->U, s->V, s->Ut, s->D, s->Dinv, s->S, s->Kinv};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete s;
}

void plr_sp_release(dplrh_plan *p)
{
    for (int i = 0; i < 2; ++i) {
        sp_free(p->sp[i]);
        p->sp[i] = nullptr;
    }
}

// type 1: FP32, 2: BF16 factor matrices.  Converts what the plan holds:
// the apply operands always, the solve operands when the FP64
// factorization is done (with `factor`, it is run first when missing).
extern "C" int dplrh_sp_factor(dplrh_plan *p, int type, int factor, int *info, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (info) *info = 0;
    if (type != 1 && type != 2) {
        plr_set_error("single-precision type %d (1: FP32, 2: BF16)", type);
        return -1;
    }
    if (factor && !p->f.done) {
        int inf = 0;
        const int rc = dplrh_factor(p, &inf, st);
        if (rc) return rc;
        if (info) *info = inf;
        if (inf) return 0;
    }
    const int bf = type == 2;
    const size_t es = bf ? sizeof(uint16_t) : sizeof(float);
    const size_t nuv = p->uoff[p->L], ns = p->soff[p->L], nd = (size_t)p->N * p->b;
    plr_sp *s = p->sp[type - 1];
    if (!s) {
        s = new plr_sp;
        memset(s, 0, sizeof(*s));
        s->type = type;
        p->sp[type - 1] = s;
        PLR_HIP(hipMalloc(&s->U, (nuv ? nuv : 1) * es));
        PLR_HIP(hipMalloc(&s->V, (nuv ? nuv : 1) * es));
        PLR_HIP(hipMalloc(&s->D, (nd ? nd : 1) * es));
        PLR_HIP(hipMalloc(&s->S, (ns ? ns : 1) * sizeof(float)));
    }
    int rc;
    if ((rc = sp_convert(p->U, nuv, bf, s->U, st)) || (rc = sp_convert(p->V, nuv, bf, s->V, st)) ||
        (rc = sp_convert(p->D, nd, bf, s->D, st)) || (rc = sp_convert(p->S, ns, 0, s->S, st)))
        return rc;
    s->fact = 0;
    if (p->f.done && !p->f.info) {
        const size_t nk = p->f.koff[p->L];
        if (!s->Ut) PLR_HIP(hipMalloc(&s->Ut, (nuv ? nuv : 1) * es));
        if (!s->Dinv) PLR_HIP(hipMalloc(&s->Dinv, (nd ? nd : 1) * es));
        if (!s->Kinv) PLR_HIP(hipMalloc(&s->Kinv, (nk ? nk : 1) * sizeof(float)));
        if ((rc = sp_convert(p->f.Ut, nuv, bf, s->Ut, st)) ||
            (rc = sp_convert(p->f.Dinv, nd, bf, s->Dinv, st)) ||
            (rc = sp_convert(p->f.Kinv, nk, 0, s->Kinv, st)))
            return rc;
        s->fact = 1;
    }
    ++s->gen;
    ++p->ws_gen; // batches rebuild their tables
    return 0;
}
