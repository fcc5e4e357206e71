// LDS read bandwidth of the device (the peak the LDS-resident sweep kernels
// are priced against).  Every block fills 64 KB of LDS, then each thread
// reads 16-byte vectors at lane-consecutive addresses (one full row of banks
// per wave instruction, no bank conflicts), iters times, and folds them into
// one value so no read can be dropped.  Bytes = blocks x threads x iters x 16.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "host.h"
#include "misamplers.h"

namespace {

constexpr int kLdsBenchThreads = 1024;
constexpr int kLdsBenchWords = 4096;  // uint4 words: 64 KB

__global__ __launch_bounds__(kLdsBenchThreads) void lds_read_kernel(int iters, uint32_t* sink) {
    __shared__ uint4 buf[kLdsBenchWords];
    for (int i = threadIdx.x; i < kLdsBenchWords; i += blockDim.x)
        buf[i] = make_uint4(i, i * 3u, i * 5u, i * 7u);
    __syncthreads();
    uint32_t acc = 0;
    int idx = threadIdx.x;
    for (int it = 0; it < iters; ++it) {
        const uint4 v = buf[idx];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
        idx = (idx + 64) & (kLdsBenchWords - 1);
    }
    if (acc == 0x9e3779b9u) sink[blockIdx.x] = acc;  // practically never; keeps the reads live
}

}  // namespace

extern "C" int misa_lds_bandwidth(int iters, double* gbs, double* ms) {
    if (iters <= 0 || !gbs) return misa_host::fail(MISA_ERR_ARG, "iters must be positive");
    int dev = 0, cus = 0;
    MISA_HIP(hipGetDevice(&dev));
    MISA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = 4 * cus;  // 16 waves per CU (LDS holds two 64 KB blocks per CU at a time)
    uint32_t* sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&sink, sizeof(uint32_t) * blocks);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(kLdsBenchThreads), 0, 0, iters, sink);  // warm-up
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(kLdsBenchThreads), 0, 0, iters, sink);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFree(sink);
    if (e != hipSuccess) return misa_host::fail(MISA_ERR_DEVICE, std::string("lds bandwidth: ") + hipGetErrorString(e));
    const double bytes = (double)blocks * kLdsBenchThreads * (double)iters * 16.0;
    *gbs = bytes / (t / 1e3) / 1e9;
    if (ms) *ms = t;
    return MISA_OK;
}
