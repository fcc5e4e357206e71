// Spin moments of sample sets: sum_r s_ri and sum_r s_ru s_rv per model.
//
// The backward pass of fixed-beta sampling: for F = -(1/beta) log Z with
// E = sum_i h_i s_i + sum_c J_c s_u s_v, dF/dh_i = <s_i> and dF/dJ_c = <s_u s_v>
// (no reference counterpart).  Samples are int8 +-1 rows (R x n per model).
// A block takes tiles of 64 reads: a wave loads one byte column of the tile
// (one lane per read) and a ballot turns it into a 64-bit mask per variable
// (bit r set when read r has s = +1), so
//   sum over the tile of s_i       = 2 popc(mask_i) - valid,
//   sum over the tile of s_u s_v   = valid - 2 popc(mask_u ^ mask_v).
// Partial sums per (model, split) are exact int32; a second kernel adds the
// splits in a fixed order and scales, so results are deterministic.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "moments_engine.h"

namespace misa {
namespace {

__global__ __launch_bounds__(256) void moments_partial_kernel(const MomentBlock* __restrict__ blocks,
                                                              const MomentModel* __restrict__ models,
                                                              const int8_t* __restrict__ samples,
                                                              const int32_t* __restrict__ edges,
                                                              int32_t* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem[];
    const MomentBlock blk = blocks[blockIdx.x];
    const MomentModel md = models[blk.model];
    const int n = md.n, m = md.m, R = md.R;
    uint64_t* sbits = reinterpret_cast<uint64_t*>(smem);
    int32_t* acc = reinterpret_cast<int32_t*>(smem + 8 * (size_t)n);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    for (int e = tid; e < n + m; e += blockDim.x) acc[e] = 0;
    const int8_t* S = samples + md.sample_off;
    const int32_t* E = edges + 2 * md.edge_off;
    __syncthreads();
    for (int tile = blk.tile_lo; tile < blk.tile_hi; ++tile) {
        const int r0 = tile * 64;
        const int valid = min(64, R - r0);
        const bool live = lane < valid;
        const int8_t* row = S + (size_t)(r0 + (live ? lane : 0)) * n;
        if ((n & 3) == 0 && ((reinterpret_cast<uintptr_t>(S) & 3) == 0)) {
            const uint32_t* row4 = reinterpret_cast<const uint32_t*>(row);
            for (int q = wave; q < (n >> 2); q += nwaves) {
                const uint32_t w = live ? row4[q] : 0x01010101u * 0xFFu;  // dead lanes read as -1
                for (int k = 0; k < 4; ++k) {
                    const uint64_t mask = __ballot(live && !((w >> (8 * k + 7)) & 1u));
                    if (lane == 0) {
                        const int i = 4 * q + k;
                        sbits[i] = mask;
                        acc[i] += 2 * __popcll(mask) - valid;
                    }
                }
            }
        } else {
            for (int i = wave; i < n; i += nwaves) {
                const int8_t v = live ? row[i] : (int8_t)-1;
                const uint64_t mask = __ballot(live && v > 0);
                if (lane == 0) {
                    sbits[i] = mask;
                    acc[i] += 2 * __popcll(mask) - valid;
                }
            }
        }
        __syncthreads();
        for (int c = tid; c < m; c += blockDim.x) {
            const int u = E[2 * c], v = E[2 * c + 1];
            acc[n + c] += valid - 2 * __popcll(sbits[u] ^ sbits[v]);
        }
        __syncthreads();
    }
    int32_t* out = part + blk.part_off;
    for (int e = tid; e < n + m; e += blockDim.x) out[e] = acc[e];
}

// grad_h[i] = scale * sum_s part[s][i] / R, grad_J[c] likewise (splits added
// in order, in int64).
__global__ __launch_bounds__(256) void moments_reduce_kernel(const MomentModel* __restrict__ models,
                                                             int num_models, const int32_t* __restrict__ part,
                                                             const float* __restrict__ scale,
                                                             float* __restrict__ grad_h, float* __restrict__ grad_J,
                                                             int64_t* __restrict__ sums) {
    const int b = blockIdx.y;
    if (b >= num_models) return;
    const MomentModel md = models[b];
    const int w = md.n + md.m;
    const float g = scale ? scale[b] : 1.0f;
    const float invR = md.R > 0 ? 1.0f / (float)md.R : 0.0f;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < w; e += gridDim.x * blockDim.x) {
        int64_t t = 0;
        for (int s = 0; s < md.splits; ++s) t += part[md.part_off + (int64_t)s * w + e];
        if (sums) sums[md.sum_off + e] = t;
        const float mean = (float)t * invR;
        if (e < md.n) {
            if (grad_h) grad_h[md.h_off + e] = g * mean;
        } else if (grad_J) {
            grad_J[md.edge_off + (e - md.n)] = g * mean;
        }
    }
}

}  // namespace

hipError_t launch_moments(const MomentBlock* d_blocks, int nblocks, const MomentModel* d_models, int num_models,
                          int max_w, int lds_bytes, const int8_t* samples, const int32_t* edges, int32_t* part,
                          const float* scale, float* grad_h, float* grad_J, int64_t* sums, hipStream_t stream) {
    if (nblocks > 0) {
        if (lds_bytes > 65536) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(moments_partial_kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(moments_partial_kernel, dim3(nblocks), dim3(256), lds_bytes, stream, d_blocks, d_models,
                           samples, edges, part);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (num_models > 0) {
        const int gx = (max_w + 255) / 256;
        hipLaunchKernelGGL(moments_reduce_kernel, dim3(gx > 0 ? gx : 1, num_models), dim3(256), 0, stream, d_models,
                           num_models, part, scale, grad_h, grad_J, sums);
        return hipGetLastError();
    }
    return hipSuccess;
}

}  // namespace misa
