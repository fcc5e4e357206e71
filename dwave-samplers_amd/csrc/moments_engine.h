// Spin-moment reduction (backward pass of fixed-beta sampling): launch
// descriptors shared by moments_api.cpp and moments_kernel.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

// One model: R samples of n int8 spins (row-major at sample_off) and m edges
// ((u, v) int32 pairs at edge_off); its splits (one per tile of 64 reads)
// partial rows of n + m int8, pw bytes apart (n + m rounded up to 16), start
// at part_off (a multiple of 16); results go to grad_h[h_off..],
// grad_J[edge_off..], sums[sum_off..].
struct MomentModel {
    int n, m, R, splits, pw, pad;
    int64_t sample_off, edge_off, h_off, part_off, sum_off;
};

// Threads of a partial block: four waves.
constexpr int kMomentThreads = 256;

// Staged loads, for rows of kStageMinVars or more variables (a multiple of
// 16): each wave passes 64 rows of 128 bytes through an LDS stage, rows
// kStageStride bytes apart.
constexpr int kStageMinVars = 64;
constexpr int kStageStride = 144;
constexpr int kStageBytes = 64 * kStageStride;
inline size_t moment_stage_lds(int n) {
    return (n % 16 == 0 && n >= kStageMinVars) ? 4 * (size_t)kStageBytes : 0;
}

// One partial block: tiles [tile_lo, tile_hi) (64 reads each) of one model;
// wpt waves build the sign masks of one tile.  At most 4 / wpt tiles, or
// kMaxTilesPerBlock when rows have at most 16 variables (wpt = 1, each wave
// takes every fourth tile).
constexpr int kMaxTilesPerBlock = 16;
// Blocks of more than four tiles keep their model's edges in LDS after the
// partial rows when there are at most kEdgeStageMax of them.
constexpr int kEdgeStageMax = 1024;
inline size_t moment_edge_lds(int ntiles, int m) {
    return ntiles > 4 && m <= kEdgeStageMax ? 8 * (size_t)m : 0;
}
// Tile t's partial row is part[part_off + (t - tile_lo) * pw ..).
// The block carries its model's descriptor, so a block's first load is its
// only dependent one.
struct MomentBlock {
    MomentModel md;
    int model, tile_lo, tile_hi, wpt;
    int64_t part_off;
};

// One reduce block: entries [e0, e0 + kReduceSpan) of one model (descriptor
// carried), four per lane.
constexpr int kReduceSpan = 256;
struct MomentSpan {
    MomentModel md;
    int model, e0;
};

// Partial sums of the given blocks (each with at most lds_bytes of LDS: the
// masks of its tiles, 8 bytes per variable rounded up to 16, then the larger
// of their partial rows, pw bytes each, and moment_stage_lds(n), which share
// one region, then moment_edge_lds(tiles, m)).
hipError_t launch_moments_partial(const MomentBlock* d_blocks, int nblocks, int lds_bytes, const int8_t* samples,
                                  const int32_t* edges, int8_t* part, hipStream_t stream);
// Per-model reduction of the partial rows, scaled: one block per span.
hipError_t launch_moments_reduce(const MomentSpan* d_spans, int nspans, const int8_t* part, const float* scale,
                                 float* grad_h, float* grad_J, int64_t* sums, hipStream_t stream);

}  // namespace misa
