// rf_solve_flow_noalloc: the TV-L1 solve for callers that cannot allocate device memory (code under
// HIP graph capture, serving loops, harnesses that forbid hipMalloc). Its working fields live in
// `scratch`, a file-scope device array of RF_NOALLOC_SCRATCH_BYTES that each GPU the library runs on
// holds from the code object's load until the process ends. Host code never names the array itself
// except to look up its address: on the first call on a device, with that device current,
// hipGetSymbolAddress gives that GPU's copy, and every later use (kernel arguments, copies) goes
// through the pointer kept for the device.
//
// Every call on a GPU shares that GPU's array and leaves its fields behind; a call zeroes the part it
// uses before its first kernel. Calls that overlap on one GPU (on different streams, or from several
// host threads without ordering) race on the array.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/rfhip.h"
#include "rf_solver.hpp"

namespace rf {

constexpr size_t kScratchFloats = RF_NOALLOC_SCRATCH_BYTES / sizeof(float);
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_scratch[kMaxDevices];  // each device's address of `scratch`, looked up once

// The current device's address of the scratch array (looked up on its first use).
int scratch_for_current_device(float** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error(std::string("rf_solve_flow_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return RF_ERR_HIP;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("rf_solve_flow_noalloc: device index out of range");
        return RF_ERR_STATE;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("rf_solve_flow_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return RF_ERR_HIP;
        }
        g_scratch[dev] = static_cast<float*>(p);
    }
    *out = g_scratch[dev];
    return RF_OK;
}
}  // namespace

}  // namespace rf

extern "C" int rf_solve_flow_noalloc(const rf_params* params, const float* d_i0, const float* d_i1, float* d_u,
                                     float* d_v, void* stream) {
    if (!params) {
        rf::set_error("rf_solve_flow_noalloc: null params");
        return RF_ERR_ARG;
    }
    float* base = nullptr;
    if (int rc = rf::scratch_for_current_device(&base)) return rc;
    rf::Solver s(32, 12, params->stride_alignment);
    return s.solve_flow_noalloc(*params, d_i0, d_i1, d_u, d_v, static_cast<hipStream_t>(stream), base,
                                rf::kScratchFloats);
}
