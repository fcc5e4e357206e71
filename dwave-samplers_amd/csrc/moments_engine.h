// Spin-moment reduction (backward pass of fixed-beta sampling): launch
// descriptors shared by moments_api.cpp and moments_kernel.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

// One model: R samples of n int8 spins (row-major at sample_off) and m edges
// ((u, v) int32 pairs at edge_off); its splits (one per tile of 64 reads)
// partial rows of n + m int8 start at part_off; results go to
// grad_h[h_off..], grad_J[edge_off..], sums[sum_off..].
struct MomentModel {
    int n, m, R, splits;
    int64_t sample_off, edge_off, h_off, part_off, sum_off;
};

// One block: tile tile_lo (64 reads) of one model, written to
// part[part_off .. part_off + n + m).
struct MomentBlock {
    int model, tile_lo, tile_hi, pad;
    int64_t part_off;
};

// Partial sums of the given blocks (each with at most lds_bytes of LDS), then
// the per-model reduction.
hipError_t launch_moments_partial(const MomentBlock* d_blocks, int nblocks, const MomentModel* d_models,
                                  int lds_bytes, int threads, const int8_t* samples, const int32_t* edges,
                                  int8_t* part, hipStream_t stream);
hipError_t launch_moments_reduce(const MomentModel* d_models, int num_models, int max_w, const int8_t* part,
                                 const float* scale, float* grad_h, float* grad_J, int64_t* sums,
                                 hipStream_t stream);

}  // namespace misa
