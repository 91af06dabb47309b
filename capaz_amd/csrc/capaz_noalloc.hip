// capaz_check_finite_noalloc: the non-finite sentinel of capaz_check_finite
// over caller-owned device arrays, for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  Its only device memory beyond the caller's arrays is `scratch`
// below, a static array of one result slot per scanned field; every call on a
// GPU shares that GPU's copy, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/capaz.h"

namespace capaz::detail {
int fail(capaz_ctx* c, int code, const std::string& msg);
}

namespace {

constexpr int kBlock = 256;
constexpr int kMaxGrid = 4096;
constexpr int kMaxDevices = 64;

// Slot q holds the smallest index at which field q is not finite (~0: none).
static __device__ __align__(256) unsigned long long scratch[CAPAZ_SENTINEL_MAX_FIELDS];

struct FieldList {
    const double* f[CAPAZ_SENTINEL_MAX_FIELDS];
};

// ~0 into the slots: a kernel rather than hipMemsetAsync, so the call may be
// captured in a graph.
__global__ void __launch_bounds__(64) k_slots_clear(unsigned long long* slots, int nfields) {
    const int q = (int)threadIdx.x;
    if (q < nfields) slots[q] = ~0ull;
}

// Grid-stride scan; per field, the lowest lane of a wavefront that sees a
// non-finite value (x fastest, so the lowest lane holds the smallest index)
// takes the atomic minimum.
__global__ void __launch_bounds__(kBlock) k_slots_scan(FieldList fl, int nfields, int64_t n,
                                                       unsigned long long* slots) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int lane = (int)(threadIdx.x & 63u);
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += stride) {
        const int64_t l = base + threadIdx.x;
        for (int q = 0; q < nfields; ++q) {
            const bool bad = l < n && !isfinite(fl.f[q][l]);
            const unsigned long long mask = __ballot(bad);
            if (mask != 0ull && lane == __ffsll((long long)mask) - 1) atomicMin(&slots[q], (unsigned long long)l);
        }
    }
}

std::mutex g_mu;
unsigned long long* g_slots[kMaxDevices] = {};

// This device's copy of `scratch`, looked up once per device with the device
// current (the array's host-side name is not a device address).
int device_slots(int dev, unsigned long long** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_slots[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return capaz::detail::fail(nullptr, CAPAZ_ERR_DEVICE,
                                       std::string("capaz_check_finite_noalloc: hipGetSymbolAddress: ") +
                                           hipGetErrorString(e));
        g_slots[dev] = static_cast<unsigned long long*>(ptr);
    }
    *out = g_slots[dev];
    return CAPAZ_OK;
}

int hip_error(hipError_t e, const char* what) {
    return capaz::detail::fail(nullptr, CAPAZ_ERR_DEVICE,
                               std::string("capaz_check_finite_noalloc: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" int capaz_check_finite_noalloc(const double* const* fields, int nfields, int64_t n, void* stream,
                                          int64_t* first_bad) {
    using capaz::detail::fail;
    if (!fields || !first_bad) return fail(nullptr, CAPAZ_ERR_INVALID, "capaz_check_finite_noalloc: NULL argument");
    if (nfields < 1 || nfields > CAPAZ_SENTINEL_MAX_FIELDS)
        return fail(nullptr, CAPAZ_ERR_INVALID,
                    "capaz_check_finite_noalloc: nfields = " + std::to_string(nfields) + " is outside 1.." +
                        std::to_string(CAPAZ_SENTINEL_MAX_FIELDS) + " (the static scratch holds one slot per field)");
    if (n < 0) return fail(nullptr, CAPAZ_ERR_INVALID, "capaz_check_finite_noalloc: n < 0");
    FieldList fl = {};
    for (int q = 0; q < nfields; ++q) {
        if (!fields[q] && n > 0)
            return fail(nullptr, CAPAZ_ERR_INVALID, "capaz_check_finite_noalloc: NULL field pointer");
        fl.f[q] = fields[q];
    }
    *first_bad = -1;
    if (n == 0) return CAPAZ_OK;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_error(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, CAPAZ_ERR_INVALID, "capaz_check_finite_noalloc: device ordinal");
    unsigned long long* slots = nullptr;
    int rc = device_slots(dev, &slots);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    const unsigned grid = (unsigned)(blocks < kMaxGrid ? blocks : kMaxGrid);
    hipLaunchKernelGGL(k_slots_clear, dim3(1), dim3(64), 0, s, slots, nfields);
    hipLaunchKernelGGL(k_slots_scan, dim3(grid), dim3(kBlock), 0, s, fl, nfields, n, slots);
    if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "launch");
    unsigned long long h[CAPAZ_SENTINEL_MAX_FIELDS];
    if ((e = hipMemcpyAsync(h, slots, (size_t)nfields * sizeof(h[0]), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return hip_error(e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_error(e, "hipStreamSynchronize");
    unsigned long long first = ~0ull;
    std::string which;
    for (int q = 0; q < nfields; ++q) first = h[q] < first ? h[q] : first;
    if (first == ~0ull) return CAPAZ_OK;
    for (int q = 0; q < nfields; ++q)
        if (h[q] == first) which += (which.empty() ? "" : ", ") + std::to_string(q);
    *first_bad = (int64_t)first;
    return fail(nullptr, CAPAZ_ERR_NONFINITE,
                "non-finite value at index " + std::to_string(first) + " in field(s) " + which);
}
