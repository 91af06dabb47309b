// Device self-tests of the kernels' arithmetic building blocks (test
// infrastructure, built by capaz_amd/build.py into tests/hip/): each compares
// a rewritten operation with the plain IEEE operation it replaces, on the
// GPU, and counts mismatching bit patterns.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "capaz_cell.h"

namespace {

__device__ uint64_t mix(uint64_t x) {  // splitmix64 finaliser
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Random doubles: uniform significand and sign, exponent in [emin, emax];
// every 16th draw has an all-ones significand, every 16th a power of two.
__device__ double draw(uint64_t seed, uint64_t i, int emin, int emax) {
    uint64_t r = mix(seed ^ mix(i));
    uint64_t mant = r & ((1ull << 52) - 1);
    uint64_t sel = (r >> 52) & 15;
    if (sel == 0) mant = (1ull << 52) - 1;
    if (sel == 1) mant = 0;
    int e = emin + (int)(mix(r) % (uint64_t)(emax - emin + 1));
    uint64_t bits = ((uint64_t)(e + 1023) << 52) | mant | ((r >> 63) << 63);
    return __longlong_as_double((long long)bits);
}

__global__ void k_rcp(uint64_t n, uint64_t seed, int emin, int emax, unsigned long long* bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (; i < n; i += stride) {
        double b = draw(seed, i, emin, emax);
        double want = 1. / b, got = capaz::rcp_rn(b);
        if (__double_as_longlong(want) != __double_as_longlong(got)) ++local;
    }
    if (local) atomicAdd(bad, local);
}

__global__ void k_div(uint64_t n, uint64_t seed, int emin, int emax, unsigned long long* bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long local = 0;
    for (; i < n; i += stride) {
        double a = draw(seed, 2 * i, emin, emax), b = draw(seed, 2 * i + 1, emin, emax);
        double want = a / b;
        double got = capaz::exact_div(a, b, capaz::rcp_rn(b));
        if (__double_as_longlong(want) != __double_as_longlong(got)) ++local;
    }
    if (local) atomicAdd(bad, local);
}

// log_cl exists on the device only (the host build of the closures calls pow)
__global__ void k_log(const double* x, double* y, uint64_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = capaz::log_cl(x[i]);
#else
    (void)x, (void)y, (void)n;
#endif
}

__global__ void k_pow(const double* x, double e, double* y, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = capaz::pow_cl(x[i], e);
}

__global__ void k_map(int which, const double* x, double* y, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = which == 0 ? capaz::pow_half(x[i]) : capaz::pow_two(x[i]);
}

}  // namespace

extern "C" {

// y = pow_half(x) (which = 0) or pow_two(x) (which = 1) on the device for n
// host doubles; 0, or -1 on a HIP error.
int capaz_selftest_map(int which, const double* x, double* y, unsigned long long n) {
    double *dx = nullptr, *dy = nullptr;
    int rc = -1;
    if (hipMalloc(&dx, n * sizeof(double)) == hipSuccess && hipMalloc(&dy, n * sizeof(double)) == hipSuccess &&
        hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(k_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, which, dx, dy, (uint64_t)n);
        if (hipGetLastError() == hipSuccess && hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

// y = log_cl(x) on the device (the closures' powers' log) for n host doubles;
// 0, or -1 on a HIP error.
int capaz_selftest_log(const double* x, double* y, unsigned long long n) {
    double *dx = nullptr, *dy = nullptr;
    int rc = -1;
    if (hipMalloc(&dx, n * sizeof(double)) == hipSuccess && hipMalloc(&dy, n * sizeof(double)) == hipSuccess &&
        hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(k_log, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx, dy, (uint64_t)n);
        if (hipGetLastError() == hipSuccess && hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

// y = pow_cl(x, e) on the device (the closures' powers) for n host doubles;
// 0, or -1 on a HIP error.
int capaz_selftest_pow(const double* x, double e, double* y, unsigned long long n) {
    double *dx = nullptr, *dy = nullptr;
    int rc = -1;
    if (hipMalloc(&dx, n * sizeof(double)) == hipSuccess && hipMalloc(&dy, n * sizeof(double)) == hipSuccess &&
        hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(k_pow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx, e, dy, (uint64_t)n);
        if (hipGetLastError() == hipSuccess && hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

// Number of draws where rcp_rn(b) != 1.0/b (which = 0) or
// exact_div(a, b, rcp_rn(b)) != a/b (which = 1); -1 on a HIP error.
long long capaz_selftest(int which, unsigned long long n, unsigned long long seed, int emin, int emax) {
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, sizeof(*d)) != hipSuccess) return -1;
    if (hipMemset(d, 0, sizeof(*d)) != hipSuccess) return -1;
    if (which == 0)
        hipLaunchKernelGGL(k_rcp, dim3(4096), dim3(256), 0, 0, (uint64_t)n, (uint64_t)seed, emin, emax, d);
    else
        hipLaunchKernelGGL(k_div, dim3(4096), dim3(256), 0, 0, (uint64_t)n, (uint64_t)seed, emin, emax, d);
    unsigned long long h = 0;
    if (hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    (void)hipFree(d);
    return (long long)h;
}

}  // extern "C"
