// Spin-moment reduction (backward pass of fixed-beta sampling): launch
// descriptors shared by moments_api.cpp and moments_kernel.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

// One model: R samples of n int8 spins (row-major at sample_off) and m edges
// ((u, v) int32 pairs at edge_off); its splits partial rows of n + m int32
// start at part_off; results go to grad_h[h_off..], grad_J[edge_off..],
// sums[sum_off..].
struct MomentModel {
    int n, m, R, splits;
    int64_t sample_off, edge_off, h_off, part_off, sum_off;
};

// One block: tiles [tile_lo, tile_hi) of 64 reads of one model, written to
// part[part_off .. part_off + n + m).
struct MomentBlock {
    int model, tile_lo, tile_hi, pad;
    int64_t part_off;
};

hipError_t launch_moments(const MomentBlock* d_blocks, int nblocks, const MomentModel* d_models, int num_models,
                          int max_w, int lds_bytes, const int8_t* samples, const int32_t* edges, int32_t* part,
                          const float* scale, float* grad_h, float* grad_J, int64_t* sums, hipStream_t stream);

}  // namespace misa
