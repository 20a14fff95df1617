// uc_probe.hip -- a plain device copy of known size, used to calibrate the
// HBM traffic counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): one launch
// reads exactly `bytes` and writes exactly `bytes`.
//
//     uc_copy_probe [MiB=1024] [launches=3]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char **argv)
{
    const size_t mib = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1024;
    const int reps = argc > 2 ? atoi(argv[2]) : 3;
    const size_t bytes = mib << 20, n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    for (int r = 0; r < reps; r++) {
        if (hipEventRecord(e0, 0) != hipSuccess) return 1;
        k_copy_probe<<<8192, 256>>>(a, b, n);
        if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
        printf("copy %zu bytes read + %zu bytes written in %.3f ms: %.1f GB/s\n", bytes, bytes, ms,
               2.0 * bytes / (ms * 1e-3) / 1e9);
    }
    if (hipFree(a) != hipSuccess || hipFree(b) != hipSuccess) return 1;
    return 0;
}
