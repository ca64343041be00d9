// Function values at points with no device allocation: the entry point
// rin_eval_points_noalloc (include/rin.h) for callers that cannot allocate
// (code under graph capture, tight serving loops).  Inputs and outputs are
// the caller's device buffers; the one intermediate, the function-major
// values that k_eval writes before they are transposed into the caller's
// row-major rows, lives in a file-scope device array of fixed size.
//
// The array's name is a host-side stand-in: its device address (one per GPU)
// comes from hipGetSymbolAddress at context creation, and only that address
// is passed to kernels.  Every call on a GPU shares that GPU's array, so
// calls that overlap on one GPU race; calls on one context's stream run in
// order.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "kernels_ia.h"
#include "pipeline.h"

namespace rin {

// function-major staging: at most kNoallocValues values (16 MiB) a call
__device__ __align__(256) double scratch[kNoallocValues];

void* noalloc_scratch_address() {
    // once per device, with that device current (the caller set it)
    static std::mutex mu;
    static std::map<int, void*> per_device;
    int dev = 0;
    RIN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_device.find(dev);
    if (it != per_device.end()) return it->second;
    void* ptr = nullptr;
    RIN_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw std::runtime_error("no device address for the allocation-free path's array");
    per_device[dev] = ptr;
    return ptr;
}

void eval_points_noalloc(void* scratch_dev, const double* pts, uint64_t n, const FuncDesc* funcs, uint32_t F,
                         double* rowmajor, hipStream_t s) {
    if (!scratch_dev) throw std::invalid_argument("no scratch address for this device");
    if (F == 0) throw std::invalid_argument("no functions");
    if (n == 0) return;
    if (n > kNoallocValues / F)
        throw std::invalid_argument("points x functions exceed the " + std::to_string(kNoallocValues) +
                                    " values of the allocation-free path");
    double* staging = static_cast<double*>(scratch_dev);
    hipLaunchKernelGGL(k_eval, dim3(nblk(n)), dim3(256), 0, s, pts, n, funcs, F, staging);
    hipLaunchKernelGGL(k_transpose_out, dim3(nblk(F * n)), dim3(256), 0, s, (const double*)staging, n, F, rowmajor);
    RIN_HIP(hipGetLastError());
}

}  // namespace rin
