// Device-side inputs of a batched run, so that a caller whose model values
// and seeds live on the GPU (differentiable fixed-beta sampling) never
// copies them through the host:
//   * values_kernel writes float32 biases into the float64 arrays a batch's
//     problems hold on the device (h, couplers in input order, the CSR
//     weights and the next-variable coupling of the pipelined loop);
//   * seeds_kernel is the engine's per-read seed derivation on the device;
//   * spins_kernel draws random +-1 initial states from a seed.
// No reference counterpart: the reference takes its inputs from the host.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sa_engine.h"

namespace misa {
namespace {

__device__ __forceinline__ uint64_t splitmix_at_d(uint64_t seed, int64_t r) {
    uint64_t z = seed + (uint64_t)(r + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Block row y = problem y; the blocks of a row stride over its arrays.
__global__ __launch_bounds__(256) void values_kernel(const ValueMap* __restrict__ maps, const float* __restrict__ h32,
                                                     const float* __restrict__ J32) {
    const ValueMap mp = maps[blockIdx.y];
    const float* hb = h32 + mp.h_off;
    const float* Jb = J32 + mp.j_off;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < mp.n; t += stride) {
        mp.h[t] = (double)hb[t];
        const int c = mp.jn_cpl[t];
        mp.jnext_d[t] = c >= 0 ? (double)Jb[c] : 0.0;
    }
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < mp.m; t += stride)
        mp.cw[t] = (double)Jb[t];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 2 * mp.m; t += stride)
        mp.Jd[t] = (double)Jb[mp.csr_cpl[t]];
}

__global__ __launch_bounds__(256) void seeds_kernel(uint64_t seed, int64_t count, uint64_t* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < count; r += stride)
        out[r] = splitmix_at_d(seed, r);
}

// Eight spins per thread and step, stored as one 8-byte word when out is
// 8-byte aligned.
__global__ __launch_bounds__(256) void spins_kernel(uint64_t seed, int64_t count, int8_t* __restrict__ out) {
    const uint64_t s = seed ^ kSpinSeedMix;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const bool vec = (reinterpret_cast<uintptr_t>(out) & 7) == 0;
    const int64_t words = vec ? count / 8 : 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < words; q += stride) {
        uint64_t wv = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            wv |= (uint64_t)((splitmix_at_d(s, 8 * q + k) >> 63) ? 0x01u : 0xFFu) << (8 * k);
        reinterpret_cast<uint64_t*>(out)[q] = wv;
    }
    for (int64_t i = 8 * words + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        out[i] = (splitmix_at_d(s, i) >> 63) ? (int8_t)1 : (int8_t)-1;
}

int grid_for(int64_t count) {
    const int64_t g = (count + 255) / 256;
    return (int)(g < 1 ? 1 : g > 4096 ? 4096 : g);
}

}  // namespace

hipError_t launch_values(const ValueMap* d_maps, int nmaps, int64_t max_len, const float* h32, const float* J32,
                         hipStream_t stream) {
    const int64_t gx = (max_len + 255) / 256;
    for (int y0 = 0; y0 < nmaps; y0 += 65535) {  // grid rows per launch
        const int ny = nmaps - y0 < 65535 ? nmaps - y0 : 65535;
        hipLaunchKernelGGL(values_kernel, dim3((unsigned)(gx < 1 ? 1 : gx > 1024 ? 1024 : gx), ny), dim3(256), 0,
                           stream, d_maps + y0, h32, J32);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_read_seeds(uint64_t seed, int64_t count, uint64_t* out, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(seeds_kernel, dim3(grid_for(count)), dim3(256), 0, stream, seed, count, out);
    return hipGetLastError();
}

hipError_t launch_random_spins(uint64_t seed, int64_t count, int8_t* out, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(spins_kernel, dim3(grid_for((count + 7) / 8)), dim3(256), 0, stream, seed, count, out);
    return hipGetLastError();
}

}  // namespace misa
