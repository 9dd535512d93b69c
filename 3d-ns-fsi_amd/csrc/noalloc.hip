/*
 * noalloc.hip — ns_totals_noalloc: the conservation totals of ns_totals on fields the caller holds in device
 * memory, for callers that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). It calls no allocating routine: its per-block partial sums live in `scratch`, a file-scope
 * device array of kTotalsBlocks x 5 doubles (10 KiB) that every GPU the library runs on holds until the
 * process ends. Its address is taken once per device with hipGetSymbolAddress (the array's name is a host
 * stand-in, not a device address); kernels receive that address.
 *
 * All calls on one GPU share that GPU's array, which keeps the last call's partials: every partial is written
 * by its block before the final kernel reads it, so nothing needs clearing. Calls that overlap on one GPU
 * (other streams or threads) race on the array and must be serialised by the caller.
 */
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/ns_fsi.h"
#include "totals.h"

namespace nsfsi {

namespace {

static __device__ __align__(256) double scratch[kTotalsBlocks * 5];

constexpr int kMaxDevices = 64;
// Largest grid accepted: 2^36 nodes (5 FP32 fields of it would already exceed the device memory).
constexpr int64_t kMaxNodes = int64_t(1) << 36;

struct ScratchTable {
    std::mutex mu;
    double* ptr[kMaxDevices] = {};
};
ScratchTable& table() {
    static ScratchTable t;
    return t;
}

// The current device's copy of `scratch` (looked up once per device, with that device current).
double* scratchOfCurrentDevice(std::string& err) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        err = "no current HIP device";
        return nullptr;
    }
    ScratchTable& t = table();
    std::lock_guard<std::mutex> lock(t.mu);
    if (!t.ptr[dev]) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            err = "the current device is not a gfx950 GPU";
            return nullptr;
        }
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        t.ptr[dev] = (double*)p;
    }
    return t.ptr[dev];
}

template <class T>
__global__ void k_totals_dense(size_t N, const T* S, const uint8_t* types, double* part) {
    totalsBlock<T>(N, [](size_t n) { return n; }, N, S, types, part);
}

// The block partials summed in block order, times the node volume (the order ns_totals sums them in).
__global__ void k_totals_final(const double* part, double vol, double* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int q = 0; q < 5; ++q) {
        double s = 0;
        for (int b = 0; b < kTotalsBlocks; ++b) s += part[(size_t)b * 5 + q];
        out[q] = s * vol;
    }
}

thread_local std::string lastError;

}  // namespace

}  // namespace nsfsi

extern "C" ns_status ns_totals_noalloc(const void* fields, int32_t precision, const uint8_t* node_types, int32_t NI,
                                       int32_t NJ, int32_t NK, const double spacing[3], double* out, void* stream) {
    using namespace nsfsi;
    if (!fields || !node_types || !spacing || !out) {
        lastError = "null pointer argument";
        return NS_ERR_ARGUMENT;
    }
    if (precision != NS_FP64 && precision != NS_FP32) {
        lastError = "precision must be NS_FP64 or NS_FP32";
        return NS_ERR_ARGUMENT;
    }
    if (NI < 1 || NJ < 1 || NK < 1) {
        lastError = "NI, NJ and NK must be positive";
        return NS_ERR_ARGUMENT;
    }
    const int64_t N = (int64_t)NI * NJ * NK;
    if (N > kMaxNodes) {
        lastError = "grid of " + std::to_string(N) + " nodes exceeds the maximum of " + std::to_string(kMaxNodes);
        return NS_ERR_ARGUMENT;
    }
    double* part = scratchOfCurrentDevice(lastError);
    if (!part) return NS_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    if (precision == NS_FP64)
        hipLaunchKernelGGL(k_totals_dense<double>, dim3(kTotalsBlocks), dim3(kTotalsThreads), 0, s, (size_t)N,
                           (const double*)fields, node_types, part);
    else
        hipLaunchKernelGGL(k_totals_dense<float>, dim3(kTotalsBlocks), dim3(kTotalsThreads), 0, s, (size_t)N,
                           (const float*)fields, node_types, part);
    const double vol = spacing[0] * spacing[1] * spacing[2];
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(64), 0, s, (const double*)part, vol, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        lastError = std::string("ns_totals_noalloc launch: ") + hipGetErrorString(e);
        return NS_ERR_HIP;
    }
    return NS_OK;
}

extern "C" const char* ns_noalloc_last_error(void) { return nsfsi::lastError.c_str(); }
