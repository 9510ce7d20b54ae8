// Calibration program for HBM traffic counters: a plain streaming copy of a
// known number of bytes.  Profiling it with FETCH_SIZE / WRITE_SIZE gives the
// factor between those counters and bytes actually moved, which is applied to
// the counters read for the SDL kernels.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void copy_f4(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 256;
    const size_t bytes = mib << 20, n = bytes / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
    for (int rep = 0; rep < 3; ++rep) hipLaunchKernelGGL(copy_f4, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    std::printf("copy_f4: %zu bytes read and %zu bytes written per launch, 3 launches\n", bytes, bytes);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
