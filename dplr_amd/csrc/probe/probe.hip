// Measurement probes for the roofline figures of this library on the GPU:
//   probe hbm   - streaming read (1 GiB) and copy (1 GiB -> 1 GiB) bandwidth;
//                 the copy kernel is also the calibration of the FETCH_SIZE /
//                 WRITE_SIZE counters (known bytes per dispatch)
//   probe mfma  - register-only throughput of v_mfma_f64_16x16x4_f64 and
//                 v_mfma_f32_16x16x32_f16
// Built by `make -C dplr_amd/csrc probe` into dplr_amd/probe.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>

typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

#define CK(x)                                                                                     \
    do {                                                                                          \
        hipError_t e = (x);                                                                       \
        if (e != hipSuccess) {                                                                    \
            printf("HIP error %s at line %d\n", hipGetErrorString(e), __LINE__);                  \
            exit(1);                                                                              \
        }                                                                                         \
    } while (0)

__global__ __launch_bounds__(256) void k_read(const d2v *__restrict__ a, size_t n2, double *out)
{
    double s = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += 4 * stride) {
        d2v v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const size_t j = i + u * stride;
            v[u] = __builtin_nontemporal_load(a + (j < n2 ? j : i));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) s += (i + u * stride < n2) ? v[u].x + v[u].y : 0.0;
    }
    if (s == 12345.678) out[0] = s;
}

__global__ __launch_bounds__(256) void k_copy(const d2v *__restrict__ a, d2v *__restrict__ b, size_t n2)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_f64(double *out, int iters)
{
    d4 acc[8];
    for (int i = 0; i < 8; ++i) acc[i] = d4{0, 0, 0, 0};
    double a = threadIdx.x * 1e-3, b = 1.0 + threadIdx.x * 1e-4;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    }
    double s = 0;
    for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    if (s == 1.2345) out[0] = s;
}

__global__ __launch_bounds__(256) void k_f16(float *out, int iters)
{
    f4 acc[8];
    for (int i = 0; i < 8; ++i) acc[i] = f4{0, 0, 0, 0};
    h8 a, b;
    for (int j = 0; j < 8; ++j) { a[j] = (_Float16)(threadIdx.x * 1e-3f); b[j] = (_Float16)1.0f; }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[i], 0, 0, 0);
    }
    float s = 0;
    for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    if (s == 1.2345f) out[0] = s;
}

static int run_hbm()
{
    const size_t bytes = (size_t)1 << 30, n2 = bytes / 16;
    d2v *a, *b;
    double *o;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMalloc(&o, 8));
    CK(hipMemset(a, 0, bytes));
    CK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    for (int grid : {1024, 2048, 4096, 8192}) {
        for (int w = 0; w < 2; ++w) hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, 0, a, n2, o);
        CK(hipEventRecord(e0));
        for (int r = 0; r < 10; ++r) hipLaunchKernelGGL(k_read, dim3(grid), dim3(256), 0, 0, a, n2, o);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        printf("read grid %d: %.1f us per GiB, %.0f GB/s\n", grid, ms * 100, bytes / (ms / 10 * 1e-3) / 1e9);
    }
    for (int grid : {2048, 8192}) {
        for (int w = 0; w < 2; ++w) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n2);
        CK(hipEventRecord(e0));
        for (int r = 0; r < 10; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, 0, a, b, n2);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        printf("copy grid %d: %.0f GB/s (read+write)\n", grid, 2.0 * bytes / (ms / 10 * 1e-3) / 1e9);
    }
    CK(hipDeviceSynchronize());
    printf("done\n");
    CK(hipFree(a));
    CK(hipFree(b));
    CK(hipFree(o));
    return 0;
}

static int run_mfma()
{
    double *o;
    CK(hipMalloc(&o, 64));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int grid = 256 * 8, iters = 2000;
    for (int k = 0; k < 2; ++k) {
        hipLaunchKernelGGL(k_f64, dim3(grid), dim3(256), 0, 0, o, 10);
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(k_f64, dim3(grid), dim3(256), 0, 0, o, iters);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double fl = (double)grid * 4 * iters * 8 * 2048.0;
        printf("f64 mfma 16x16x4: %.1f TFLOP/s (%.3f ms)\n", fl / (ms * 1e-3) / 1e12, ms);
    }
    for (int k = 0; k < 2; ++k) {
        hipLaunchKernelGGL(k_f16, dim3(grid), dim3(256), 0, 0, (float *)o, 10);
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(k_f16, dim3(grid), dim3(256), 0, 0, (float *)o, iters);
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double fl = (double)grid * 4 * iters * 8 * (16.0 * 16 * 32 * 2);
        printf("f16 mfma 16x16x32: %.1f TFLOP/s (%.3f ms)\n", fl / (ms * 1e-3) / 1e12, ms);
    }
    CK(hipDeviceSynchronize());
    printf("done\n");
    CK(hipFree(o));
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && strcmp(argv[1], "hbm") == 0) return run_hbm();
    if (argc > 1 && strcmp(argv[1], "mfma") == 0) return run_mfma();
    printf("usage: probe hbm|mfma\n");
    return 2;
}
