// hbm_copy_probe.hip -- calibration for HBM byte counters.
//
// Copies a buffer of known size once with a plain grid-stride kernel, so
// that a profiler's FETCH_SIZE / WRITE_SIZE readings for this dispatch can be
// compared with the bytes the kernel must move (argv[1] MiB read and the same
// written).  The ratio scales the same counters read on the separation
// kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy_kernel(const double4* __restrict__ src, double4* __restrict__ dst, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv)
{
    const long mib = argc > 1 ? std::atol(argv[1]) : 256;
    if (mib < 1 || mib > 4096) {
        std::fprintf(stderr, "size must be 1-4096 MiB\n");
        return 2;
    }
    const long bytes = mib << 20;
    const long n = bytes / (long)sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) {
        std::fprintf(stderr, "copy failed\n");
        return 1;
    }
    std::printf("copied %ld bytes (%ld KiB read, %ld KiB written)\n", bytes, bytes >> 10, bytes >> 10);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
