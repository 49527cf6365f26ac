// HBM probe: a plain streaming copy of a known number of bytes, used to
// calibrate the rocprofv3 FETCH_SIZE / WRITE_SIZE counters (bytes reported
// per byte moved) and to measure the achievable copy and read bandwidth of
// the device the decoder runs on.
//
//   hbm_probe [MiB]      copies MiB (default 1024) from one buffer to another
//                        10 times, then reads it 10 times, and prints GB/s.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define PROBE_CHECK(x)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
            return 1;                                                                     \
        }                                                                                 \
    } while (0)

__global__ void probe_copy_kernel(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void probe_read_kernel(const float4 *__restrict__ a, float *__restrict__ out, size_t n) {
    float s = 0.0f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == 1234.5f) out[0] = s;  // keeps the loads live; never true for zeroed input
}

int main(int argc, char **argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1024;
    const size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    float *o = nullptr;
    PROBE_CHECK(hipMalloc(&a, bytes));
    PROBE_CHECK(hipMalloc(&b, bytes));
    PROBE_CHECK(hipMalloc(&o, sizeof(float)));
    PROBE_CHECK(hipMemset(a, 0, bytes));
    PROBE_CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    PROBE_CHECK(hipEventCreate(&e0));
    PROBE_CHECK(hipEventCreate(&e1));
    const int reps = 10;
    float ms = 0.0f;
    hipLaunchKernelGGL(probe_copy_kernel, dim3(2048), dim3(256), 0, 0, a, b, n);
    PROBE_CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(probe_copy_kernel, dim3(2048), dim3(256), 0, 0, a, b, n);
    PROBE_CHECK(hipEventRecord(e1));
    PROBE_CHECK(hipEventSynchronize(e1));
    PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("copy %zu MiB: %.1f GB/s (read+write), %.1f us per launch\n", mib,
                2.0 * bytes * reps / (ms * 1e-3) / 1e9, ms * 1e3 / reps);
    PROBE_CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(probe_read_kernel, dim3(2048), dim3(256), 0, 0, a, o, n);
    PROBE_CHECK(hipEventRecord(e1));
    PROBE_CHECK(hipEventSynchronize(e1));
    PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::printf("read %zu MiB: %.1f GB/s, %.1f us per launch\n", mib, (double)bytes * reps / (ms * 1e-3) / 1e9,
                ms * 1e3 / reps);
    PROBE_CHECK(hipFree(a));
    PROBE_CHECK(hipFree(b));
    PROBE_CHECK(hipFree(o));
    return 0;
}
