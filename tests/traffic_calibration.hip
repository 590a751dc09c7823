// Calibration of the rocprofv3 FETCH_SIZE / WRITE_SIZE counters: one kernel
// that reads 1 GiB and writes 1 GiB (16-byte loads and stores, grid-stride).
// Build: hipcc --offload-arch=gfx950 -O3 tests/traffic_calibration.hip -o <exe>
#include <hip/hip_runtime.h>
#include <stdio.h>

__global__ void k_copy_1gib(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

int main()
{
    const size_t bytes = 1ULL << 30, n = bytes / sizeof(uint4);
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess) return 1;
    hipLaunchKernelGGL(k_copy_1gib, dim3(8192), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    uint4 h;
    if (hipMemcpy(&h, b + n - 1, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 1;
    printf("copied %zu bytes, last word %u\n", bytes, h.x);
    hipFree(a);
    hipFree(b);
    return 0;
}
