// copy_probe -- device-to-device copies of known size (1 GiB read + 1 GiB written per launch).
// copy_kernel: a plain grid-stride copy; rocprofv3's FETCH_SIZE / WRITE_SIZE for it calibrate those
// counters to bytes. copy_nt_kernel: one 16-byte element per lane of a grid covering the buffer
// once, with nontemporal loads and stores -- the fastest copy found (a reference point for
// streaming kernels, not a ceiling). Both are event-timed; one JSON line each.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

#define CHK(x)                                                                      \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));     \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

// Grid-stride copy, four independent 16-byte loads in flight per thread.
__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    constexpr int U = 4;
    const size_t bd = blockDim.x;
    const size_t stride = (size_t)gridDim.x * bd * U;
    size_t i = blockIdx.x * bd * U + threadIdx.x;
    for (; i + (U - 1) * bd < n; i += stride) {
        float4 t[U];
#pragma unroll
        for (int u = 0; u < U; ++u) t[u] = a[i + u * bd];
#pragma unroll
        for (int u = 0; u < U; ++u) b[i + u * bd] = t[u];
    }
    for (; i < n; i += bd) b[i] = a[i];
}

// One grid pass over the buffer, four 16-byte elements per lane, nontemporal hints.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) copy_nt_kernel(const v4f* __restrict__ a, v4f* __restrict__ b, size_t n) {
    constexpr int U = 4;
    const size_t i = ((size_t)blockIdx.x * U) * blockDim.x + threadIdx.x;
    v4f t[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (i + (size_t)u * blockDim.x < n) t[u] = __builtin_nontemporal_load(&a[i + (size_t)u * blockDim.x]);
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (i + (size_t)u * blockDim.x < n) __builtin_nontemporal_store(t[u], &b[i + (size_t)u * blockDim.x]);
}

int main(int argc, char** argv) {
    const size_t bytes = (size_t)1 << 30;
    const size_t n = bytes / sizeof(float4);
    const int iters = argc > 1 ? std::atoi(argv[1]) : 20;
    float4 *a = nullptr, *b = nullptr;
    CHK(hipMalloc(&a, bytes));
    CHK(hipMalloc(&b, bytes));
    CHK(hipMemset(a, 0, bytes));
    CHK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    const int blocks = 256 * 16;   // 16 workgroups per CU, grid-stride over the buffer
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0));
    for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
    CHK(hipGetLastError());
    CHK(hipEventRecord(e1));
    CHK(hipEventSynchronize(e1));
    float ms = 0;
    CHK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("{\"kernel\": \"copy_kernel\", \"bytes_per_launch\": %zu, \"avg_us\": %.2f, \"GBps\": %.1f}\n",
                2 * bytes, ms * 1e3 / iters, 2.0 * bytes * iters / (ms * 1e-3) / 1e9);
    const unsigned nt_blocks = (unsigned)((n + 4 * 256 - 1) / (4 * 256));
    const v4f* av = reinterpret_cast<const v4f*>(a);
    v4f* bv = reinterpret_cast<v4f*>(b);
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(copy_nt_kernel, dim3(nt_blocks), dim3(256), 0, 0, av, bv, n);
    CHK(hipGetLastError());
    CHK(hipDeviceSynchronize());
    CHK(hipEventRecord(e0));
    for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(copy_nt_kernel, dim3(nt_blocks), dim3(256), 0, 0, av, bv, n);
    CHK(hipGetLastError());
    CHK(hipEventRecord(e1));
    CHK(hipEventSynchronize(e1));
    CHK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("{\"kernel\": \"copy_nt_kernel\", \"bytes_per_launch\": %zu, \"avg_us\": %.2f, \"GBps\": %.1f}\n",
                2 * bytes, ms * 1e3 / iters, 2.0 * bytes * iters / (ms * 1e-3) / 1e9);
    CHK(hipFree(a));
    CHK(hipFree(b));
    return 0;
}
