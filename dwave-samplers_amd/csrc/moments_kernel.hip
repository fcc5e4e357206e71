// Spin moments of sample sets: sum_r s_ri and sum_r s_ru s_rv per model.
//
// The backward pass of fixed-beta sampling: for F = -(1/beta) log Z with
// E = sum_i h_i s_i + sum_c J_c s_u s_v, dF/dh_i = <s_i> and dF/dJ_c = <s_u s_v>
// (no reference counterpart).  Samples are int8 +-1 rows (R x n per model).
// For a tile of 64 reads, a 64-bit mask per variable (bit r set when read r
// has s = +1) gives
//   sum over the tile of s_i       = 2 popc(mask_i) - valid,
//   sum over the tile of s_u s_v   = valid - 2 popc(mask_u ^ mask_v).
// Tile sums are exact int8 partials; a second kernel adds the tiles of each
// model in a fixed order and scales, so results are deterministic.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "moments_engine.h"

namespace misa {
namespace {

// One block per tile of 64 reads.  The tile's rows are copied into LDS with
// coalesced loads (row stride n + 4 bytes keeps a column read free of bank
// conflicts); each wave then reads one 4-byte column (one lane per read) and
// four ballots give the 64-bit masks of four variables.  The tile's sums lie
// in [-64, 64] and are written as int8 partials.
__global__ __launch_bounds__(512) void moments_partial_kernel(const MomentBlock* __restrict__ blocks,
                                                              const MomentModel* __restrict__ models,
                                                              const int8_t* __restrict__ samples,
                                                              const int32_t* __restrict__ edges,
                                                              int8_t* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem[];
    const MomentBlock blk = blocks[blockIdx.x];
    const MomentModel md = models[blk.model];
    const int n = md.n, m = md.m, R = md.R;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int r0 = blk.tile_lo * 64;
    const int valid = min(64, R - r0);
    const int n4 = (n + 3) >> 2;          // dwords per staged row
    const int stride = n4 + 1;             // padded row stride in dwords
    uint32_t* tile = reinterpret_cast<uint32_t*>(smem);
    uint64_t* sbits = reinterpret_cast<uint64_t*>(smem + 4 * (size_t)stride * 64 + 8);
    sbits = reinterpret_cast<uint64_t*>((reinterpret_cast<uintptr_t>(sbits) + 7) & ~(uintptr_t)7);
    const int8_t* S = samples + md.sample_off + (size_t)r0 * n;
    const int32_t* E = edges + 2 * md.edge_off;
    // Stage: byte b of row r goes to byte b of dword row r; dead rows and the
    // tail of a row are -1 (bit clear, contributes nothing: masked by valid).
    const int total = valid * n;
    if ((n & 3) == 0 && ((reinterpret_cast<uintptr_t>(S) & 3) == 0)) {
        const uint32_t* S4 = reinterpret_cast<const uint32_t*>(S);
        if (n4 >= 64) {
            // threads = column x row-group; 16 rows of loads in flight each
            const int cols = min(n4, (int)blockDim.x);
            const int groups = blockDim.x / cols;
            const int g = tid / cols, qq = tid - g * cols;
            if (g < groups)
                for (int q = qq; q < n4; q += cols)
                    for (int rb = 16 * g; rb < 64; rb += 16 * groups) {
                        uint32_t v[16];
#pragma unroll
                        for (int k = 0; k < 16; ++k) v[k] = rb + k < valid ? S4[(rb + k) * n4 + q] : 0xFFFFFFFFu;
#pragma unroll
                        for (int k = 0; k < 16; ++k) tile[(rb + k) * stride + q] = v[k];
                    }
        } else {
            for (int t = tid; t < (total >> 2); t += blockDim.x) {
                const int r = t / n4, q = t - r * n4;
                tile[r * stride + q] = S4[t];
            }
        }
    } else {
        for (int t = tid; t < 64 * n4; t += blockDim.x) {
            const int r = t / n4, q = t - r * n4;
            uint32_t w = 0xFFFFFFFFu;
            if (r < valid)
                for (int k = 0; k < 4; ++k) {
                    const int i = 4 * q + k;
                    if (i < n) w = (w & ~(0xFFu << (8 * k))) | ((uint32_t)(uint8_t)S[(size_t)r * n + i] << (8 * k));
                }
            tile[r * stride + q] = w;
        }
    }
    __syncthreads();
    const bool live = lane < valid;
    for (int q0 = wave * 4; q0 < n4; q0 += nwaves * 4) {
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = (live && q0 + j < n4) ? tile[lane * stride + q0 + j] : 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t mask = __ballot(live && !((w[j] >> (8 * k + 7)) & 1u));
                const int i = 4 * (q0 + j) + k;
                if (lane == 0 && i < n) sbits[i] = mask;
            }
    }
    __syncthreads();
    int8_t* out = part + blk.part_off;
    for (int i = tid; i < n; i += blockDim.x) out[i] = (int8_t)(2 * __popcll(sbits[i]) - valid);
    const int2* E2 = reinterpret_cast<const int2*>(E);
    for (int c0 = tid; c0 < m; c0 += 8 * blockDim.x) {
        int2 e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = c0 + j * blockDim.x;
            e[j] = c < m ? E2[c] : make_int2(0, 0);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = c0 + j * blockDim.x;
            if (c < m) out[n + c] = (int8_t)(valid - 2 * __popcll(sbits[e[j].x] ^ sbits[e[j].y]));
        }
    }
}

// grad_h[i] = scale * sum_s part[s][i] / R, grad_J[c] likewise (splits added
// in order, in int64).
__global__ __launch_bounds__(256) void moments_reduce_kernel(const MomentModel* __restrict__ models,
                                                             int num_models, const int8_t* __restrict__ part,
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
        const int8_t* p = part + md.part_off + e;
        int s = 0;
        for (; s + 16 <= md.splits; s += 16) {
            int v[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = p[(int64_t)(s + k) * w];
#pragma unroll
            for (int k = 0; k < 16; ++k) t += v[k];
        }
        for (; s < md.splits; ++s) t += p[(int64_t)s * w];
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

hipError_t launch_moments_partial(const MomentBlock* d_blocks, int nblocks, const MomentModel* d_models,
                                  int lds_bytes, int threads, const int8_t* samples, const int32_t* edges,
                                  int8_t* part, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    if (lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(moments_partial_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(moments_partial_kernel, dim3(nblocks), dim3(threads), lds_bytes, stream, d_blocks, d_models,
                       samples, edges, part);
    return hipGetLastError();
}

hipError_t launch_moments_reduce(const MomentModel* d_models, int num_models, int max_w, const int8_t* part,
                                 const float* scale, float* grad_h, float* grad_J, int64_t* sums,
                                 hipStream_t stream) {
    if (num_models <= 0) return hipSuccess;
    const int gx = (max_w + 255) / 256;
    hipLaunchKernelGGL(moments_reduce_kernel, dim3(gx > 0 ? gx : 1, num_models), dim3(256), 0, stream, d_models,
                       num_models, part, scale, grad_h, grad_J, sums);
    return hipGetLastError();
}

}  // namespace misa
