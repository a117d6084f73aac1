// copy_calib -- a plain device copy of a known size, for scaling HBM byte counters.
//
// Reads N MiB and writes N MiB per launch (16-byte loads and stores, grid-stride,
// every access inside the two buffers), three launches.  A counter run over this
// program reads FETCH_SIZE and WRITE_SIZE for a kernel whose true traffic is
// known, and the ratio scales the same counters read on the update kernel.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const double2* __restrict__ x, double2* __restrict__ y, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = x[q];
}

static int check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "copy_calib: %s: %s\n", what, hipGetErrorString(e));
        return 1;
    }
    return 0;
}

int main(int argc, char** argv)
{
    const long mib = argc > 1 ? std::atol(argv[1]) : 1024;
    if (mib <= 0 || mib > 16384) {
        std::fprintf(stderr, "copy_calib: size must be 1..16384 MiB\n");
        return 2;
    }
    const size_t bytes = (size_t)mib << 20;
    const size_t n = bytes / sizeof(double2);
    double2 *x = nullptr, *y = nullptr;
    if (check(hipMalloc(&x, bytes), "hipMalloc") || check(hipMalloc(&y, bytes), "hipMalloc")) return 1;
    if (check(hipMemset(x, 0, bytes), "hipMemset")) return 1;
    const int threads = 256;
    const int blocks = 256 * 8 * 4;
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(threads), 0, 0, x, y, n);
    if (check(hipGetLastError(), "launch") || check(hipDeviceSynchronize(), "sync")) return 1;
    std::printf("copy_calib: %zu bytes read and %zu bytes written per launch, 3 launches\n", bytes, bytes);
    hipFree(x);
    hipFree(y);
    return 0;
}
