// Allocation-free device run (misa_sa_run_device_noalloc, include/misamplers.h).
//
// The only device memory misa_sa_run_device may allocate is the general
// kernel's per-read state (fields and spins of reads whose state does not fit
// LDS, sa_sweep_kernel), which it grows on the plan.  This entry point keeps
// that state in `scratch`, a file-scope device array of kScratchBytes that
// every GPU the library runs on holds until the process ends; everything else
// a run needs (tables, run cuts, adjacency) was made when the plan was
// created.  The array's name is a host-side stand-in in host code: its device
// address comes from hipGetSymbolAddress, once per device, and only that
// address is passed to kernels.  Calls that overlap on one GPU share the array
// and race; the kernels write every byte of it they read, so nothing needs
// clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "host.h"
#include "misamplers.h"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once per device).
int scratch_address(void** out) {
    int dev = 0;
    MISA_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return misa_host::fail(MISA_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MISA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return misa_host::fail(MISA_ERR_DEVICE, "no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MISA_OK;
}

}  // namespace

extern "C" {

int misa_sa_run_device_noalloc(misa_sa_plan* plan, int8_t* d_states, double* d_energies, const uint64_t* d_seeds,
                               int num_samples, uint64_t* d_counters, void* stream) {
    void* ptr = nullptr;
    int rc = scratch_address(&ptr);
    if (rc) return rc;
    return misa_host::sa_run_device_impl(plan, d_states, d_energies, d_seeds, num_samples, d_counters, stream, ptr,
                                         kScratchBytes);
}

}  // extern "C"
