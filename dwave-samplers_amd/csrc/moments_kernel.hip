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
// model in integers and scales, so results do not depend on the launch.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "moments_engine.h"

namespace misa {
namespace {

// A block of four waves takes 4 / wpt tiles of one model; wpt waves build one
// tile's masks, each over its share of the 16-variable chunks of the rows.
// Lane r holds 64 bytes of read r's row (four chunks; rows of 64 or more
// variables are first staged through LDS with eight lanes loading one row's
// 128-byte line, so that a load touches eight cache lines instead of 64); the
// sign bits of its 64 variables form a 64-bit row of a 64 x 64 bit matrix
// (reads x variables), and a transpose across the wave leaves lane c with the
// mask of one variable, which it stores to LDS.  A thread's first kEdgePre
// edges are loaded before the masks are built (unconditional loads of clamped
// indexes, so they are all in flight at once), later ones a batch ahead of
// their use; blocks of many small tiles keep their model's edges in LDS.
// After the barrier the block's threads write one int8 partial per tile and
// entry (variables, then edges; each edge loaded once for all the block's
// tiles) into LDS rows, and after a second barrier the rows go out as 16-byte
// stores.  Samples are +-1: a clear sign bit is +1.
constexpr int kEdgePre = 16;   // edges per thread loaded before the masks
constexpr int kEdgeBatch = 8;  // edges per thread per later batch

// x of lane (lane ^ K), for K < 32.
template <int K>
__device__ __forceinline__ uint32_t xor_lane(uint32_t x) {
    if constexpr (K == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    if constexpr (K == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    if constexpr (K == 8) return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false); // row_ror:8
    if constexpr (K == 4 || K == 16) return __builtin_amdgcn_ds_swizzle(x, 0x1F | (K << 10));  // xor K
}

// One step of the transpose of the 32 x 32 bit blocks held by 32 lanes (row
// = lane, column = bit): lanes with bit K of the lane index clear exchange
// their columns with bit K set against the partner's (lane ^ K) columns with
// bit K clear.
template <int K>
__device__ __forceinline__ void transpose_step(uint32_t& lo, uint32_t& hi, int lane) {
    constexpr uint32_t m = K == 16 ? 0x0000FFFFu : K == 8 ? 0x00FF00FFu : K == 4 ? 0x0F0F0F0Fu
                         : K == 2 ? 0x33333333u : 0x55555555u;  // columns with bit K clear
    const bool clear = !(lane & K);
    const uint32_t keep = clear ? m : ~m;
    const uint32_t rot = clear ? 32 - K : K;  // partner's columns moved up (clear) or down (set) by K
    const uint32_t pl = xor_lane<K>(lo), ph = xor_lane<K>(hi);
    lo = (keep & lo) | (~keep & __builtin_amdgcn_alignbit(pl, pl, rot));
    hi = (keep & hi) | (~keep & __builtin_amdgcn_alignbit(ph, ph, rot));
}

// 64 x 64 bit transpose across the wave: lane r holds row r (bit c of
// hi:lo = column c); on return lane c holds column c (bit r = row r).  The
// off-diagonal 32 x 32 blocks are swapped (v_permlane32_swap), then each
// block is transposed in five steps.
__device__ __forceinline__ void transpose64(uint32_t& lo, uint32_t& hi) {
    const int lane = threadIdx.x & 63;
    const auto sw = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
    lo = sw[0];
    hi = sw[1];
    transpose_step<16>(lo, hi, lane);
    transpose_step<8>(lo, hi, lane);
    transpose_step<4>(lo, hi, lane);
    transpose_step<2>(lo, hi, lane);
    transpose_step<1>(lo, hi, lane);
}

// Masks of 64 variables of the wave's 64 reads: lane r holds d[0..15], the
// int8 spins of read r (variable 4 j + b in byte b of d[j]).  The sign bits
// are gathered into a 64-bit row (bit 8 b + j of the low word for j < 8, of
// the high word for j >= 8) and transposed; on return lane c holds the mask
// of variable mask_var(c) (bit r set when read r has s = +1).
__device__ __forceinline__ uint64_t masks64(const uint32_t (&d)[16]) {
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        lo |= (d[j] >> (7 - j)) & (0x01010101u << j);
        hi |= (d[8 + j] >> (7 - j)) & (0x01010101u << j);
    }
    lo = ~lo;  // set bit = clear sign bit = +1
    hi = ~hi;
    transpose64(lo, hi);
    return (uint64_t)hi << 32 | lo;
}

// The variable (0..63) whose mask lane c holds after masks64.
__device__ __forceinline__ int mask_var(int c) { return 32 * (c >> 5) + 4 * (c & 7) + ((c >> 3) & 3); }

// Masks of 16 variables into lanes 0..15: pred(b) is the lane's predicate for
// variable b, and lane b keeps ballot(pred(b)).
template <typename Pred>
__device__ __forceinline__ uint64_t chunk_masks(Pred pred) {
    const int lane = threadIdx.x & 63;
    uint64_t mine = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const uint64_t mask = __ballot(pred(b));
        mine = lane == b ? mask : mine;
    }
    return mine;
}

__global__ __launch_bounds__(kMomentThreads) void moments_partial_kernel(const MomentBlock* __restrict__ blocks,
                                                                          const int8_t* __restrict__ samples,
                                                                          const int32_t* __restrict__ edges,
                                                                          int8_t* __restrict__ part) {
    extern __shared__ __align__(16) uint64_t smask[];
    const MomentBlock blk = blocks[blockIdx.x];
    const MomentModel& md = blk.md;
    const int n = md.n, m = md.m, R = md.R, pw = md.pw;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wpt = blk.wpt;
    const int ntiles = blk.tile_hi - blk.tile_lo;
    // partial rows in LDS after the masks of all the block's tiles
    int8_t* rows = reinterpret_cast<int8_t*>(smask + (((size_t)ntiles * n + 1) & ~(size_t)1));
    const int2* E2 = reinterpret_cast<const int2*>(edges) + md.edge_off;
    int2 pre[kEdgePre];
#pragma unroll
    for (int j = 0; j < kEdgePre; ++j) {
        pre[j] = make_int2(0, 0);
        if (j * kMomentThreads < m) pre[j] = E2[min(tid + j * kMomentThreads, m - 1)];
    }
    const int8_t* base = samples + md.sample_off;
    const int nch = (n + 15) >> 4;
    const bool aligned = (n & 15) == 0 && (reinterpret_cast<uintptr_t>(base) & 15) == 0;
    static_assert(kMaxTilesPerBlock == 16, "four 16-variable tiles per wave make one 64-variable transpose");
    if (nch == 1 && wpt == 1 && aligned) {
        // one 16-byte row per read (n = 16): the wave's tiles w, w + 4, w + 8,
        // w + 12 load together and make one 64-variable transpose
        uint32_t d[16];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = wave + 4 * k;
            const int r = (blk.tile_lo + t) * 64 + lane;
            uint4 v = make_uint4(~0u, ~0u, ~0u, ~0u);
            if (t < ntiles && r < R) v = *reinterpret_cast<const uint4*>(base + (size_t)r * 16);
            d[4 * k] = v.x;
            d[4 * k + 1] = v.y;
            d[4 * k + 2] = v.z;
            d[4 * k + 3] = v.w;
        }
        const uint64_t mine = masks64(d);
        const int c = mask_var(lane);
        const int t = wave + 4 * (c >> 4);
        if (t < ntiles) smask[(size_t)t * 16 + (c & 15)] = mine;
    }
    for (int u = wave; u < ntiles * wpt && !(nch == 1 && wpt == 1 && aligned); u += 4) {
        const int t = u / wpt, p = u - t * wpt;
        const int r0 = (blk.tile_lo + t) * 64;
        const bool live = r0 + lane < R;
        const int8_t* row = base + (size_t)(live ? r0 + lane : r0) * n;  // dead lanes load nothing
        uint64_t* M = smask + (size_t)t * n;
        const int c_lo = (int)((int64_t)p * nch / wpt), c_hi = (int)((int64_t)(p + 1) * nch / wpt);
        if (aligned && n >= kStageMinVars) {
            // Rows at least kStageMinVars bytes apart: a lane per read would
            // touch one cache line per lane, so eight lanes load a row's 128
            // bytes together and the wave's rows pass through its LDS stage.
            uint8_t* stage = reinterpret_cast<uint8_t*>(rows) + (size_t)wave * kStageBytes;  // rows come later
            const int seg = lane & 7, rsub = lane >> 3;
            auto load = [&](int c0, uint4* g) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int r = 8 * i + rsub, ch = c0 + seg;
                    g[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
                    if (r0 + r < R && ch < c_hi)
                        g[i] = *reinterpret_cast<const uint4*>(base + (size_t)(r0 + r) * n + 16 * ch);
                }
            };
            uint4 g[8];
            load(c_lo, g);
            for (int c0 = c_lo; c0 < c_hi; c0 += 8) {
                __builtin_amdgcn_wave_barrier();  // the previous batch's reads stay before these writes
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    *reinterpret_cast<uint4*>(stage + (8 * i + rsub) * kStageStride + 16 * seg) = g[i];
                __builtin_amdgcn_wave_barrier();
                uint4 gn[8];
                if (c0 + 8 < c_hi) load(c0 + 8, gn);
                // two 64-variable transposes over the stage's chunks
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int cq = c0 + 4 * q;
                    if (cq < c_hi) {
                        uint32_t d[16];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const uint4 v = *reinterpret_cast<const uint4*>(stage + lane * kStageStride + 16 * (4 * q + k));
                            d[4 * k] = v.x;
                            d[4 * k + 1] = v.y;
                            d[4 * k + 2] = v.z;
                            d[4 * k + 3] = v.w;
                        }
                        const uint64_t mine = masks64(d);
                        const int c = mask_var(lane);
                        if (cq + (c >> 4) < c_hi) M[16 * cq + c] = mine;
                    }
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) g[i] = gn[i];
            }
        } else if (aligned) {
            for (int c0 = c_lo; c0 < c_hi; c0 += 8) {
                uint4 v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    v[k] = make_uint4(~0u, ~0u, ~0u, ~0u);
                    if (live && c0 + k < c_hi) v[k] = *reinterpret_cast<const uint4*>(row + 16 * (c0 + k));
                }
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int cq = c0 + 4 * q;
                    if (cq < c_hi) {
                        uint32_t d[16];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            d[4 * k] = v[4 * q + k].x;
                            d[4 * k + 1] = v[4 * q + k].y;
                            d[4 * k + 2] = v[4 * q + k].z;
                            d[4 * k + 3] = v[4 * q + k].w;
                        }
                        const uint64_t mine = masks64(d);
                        const int c = mask_var(lane);
                        if (cq + (c >> 4) < c_hi) M[16 * cq + c] = mine;
                    }
                }
            }
        } else {
            for (int c = c_lo; c < c_hi; ++c) {
                const uint64_t mine = chunk_masks([&](int b) {
                    const int i = 16 * c + b;
                    return live && i < n && row[i] >= 0;
                });
                if (lane < 16 && 16 * c + lane < n) M[16 * c + lane] = mine;
            }
        }
    }
    // blocks of many small tiles read their model's edges from LDS, after the
    // partial rows (each thread's first edges are its share)
    static_assert(kEdgeStageMax <= kEdgePre * kMomentThreads, "staged edges come from the preloaded ones");
    int2* Es = reinterpret_cast<int2*>(rows + (size_t)ntiles * pw);
    const bool edges_lds = ntiles > 4 && m <= kEdgeStageMax;
    if (edges_lds) {
#pragma unroll
        for (int j = 0; j < kEdgePre; ++j)
            if (tid + j * kMomentThreads < m) Es[tid + j * kMomentThreads] = pre[j];
    }
    __syncthreads();
    // Partials into the LDS rows, reads batched ahead of their writes (the
    // rows and masks share LDS, so a write orders every later read).
    if (ntiles > 4) {
        // many small tiles: a thread takes an entry (a variable, or an edge
        // read once) across the block's tiles, four tiles at a time
        const int w = n + m;
        for (int e = tid; e < w; e += kMomentThreads) {
            const bool var = e < n;
            int2 uv = make_int2(e, e);
            if (!var) uv = edges_lds ? Es[e - n] : E2[e - n];
            for (int t0 = 0; t0 < ntiles; t0 += 4) {
                uint64_t a[4], b[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint64_t* M = smask + (size_t)min(t0 + k, ntiles - 1) * n;
                    a[k] = M[uv.x];
                    b[k] = var ? 0ull : M[uv.y];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int tt = t0 + k;
                    if (tt < ntiles) {
                        const int val = min(64, R - (blk.tile_lo + tt) * 64);
                        const int sv = var ? 2 * __popcll(a[k]) - val : val - 2 * __popcll(a[k] ^ b[k]);
                        rows[(size_t)tt * pw + e] = (int8_t)sv;
                    }
                }
            }
        }
    }
    for (int tt = 0; tt < ntiles && ntiles <= 4; ++tt) {
        const int valid = min(64, R - (blk.tile_lo + tt) * 64);
        const uint64_t* M = smask + (size_t)tt * n;
        int8_t* rw = rows + (size_t)tt * pw;
        for (int e0 = tid; e0 < n; e0 += 4 * kMomentThreads) {
            uint64_t a[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = e0 + k * kMomentThreads;
                a[k] = e < n ? M[e] : 0ull;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = e0 + k * kMomentThreads;
                if (e < n) rw[e] = (int8_t)(2 * __popcll(a[k]) - valid);
            }
        }
#pragma unroll
        for (int j0 = 0; j0 < kEdgePre; j0 += 4) {
            if (j0 * kMomentThreads < m) {
                uint64_t a[4], b[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    a[k] = M[pre[j0 + k].x];
                    b[k] = M[pre[j0 + k].y];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int c = tid + (j0 + k) * kMomentThreads;
                    if (c < m) rw[n + c] = (int8_t)(valid - 2 * __popcll(a[k] ^ b[k]));
                }
            }
        }
    }
    // edges beyond the first kEdgePre per thread, a batch ahead
    const int c_rest = tid + kEdgePre * kMomentThreads;
    if (ntiles <= 4 && kEdgePre * kMomentThreads < m) {
        int2 cur[kEdgeBatch];
#pragma unroll
        for (int j = 0; j < kEdgeBatch; ++j) cur[j] = E2[min(c_rest + j * kMomentThreads, m - 1)];
        for (int c0 = c_rest; c0 - tid < m; c0 += kEdgeBatch * kMomentThreads) {
            int2 nxt[kEdgeBatch];
#pragma unroll
            for (int j = 0; j < kEdgeBatch; ++j) nxt[j] = E2[min(c0 + (kEdgeBatch + j) * kMomentThreads, m - 1)];
            for (int tt = 0; tt < ntiles; ++tt) {
                const int valid = min(64, R - (blk.tile_lo + tt) * 64);
                const uint64_t* M = smask + (size_t)tt * n;
                uint64_t a[kEdgeBatch], b[kEdgeBatch];
#pragma unroll
                for (int j = 0; j < kEdgeBatch; ++j) {
                    a[j] = M[cur[j].x];
                    b[j] = M[cur[j].y];
                }
#pragma unroll
                for (int j = 0; j < kEdgeBatch; ++j) {
                    const int c = c0 + j * kMomentThreads;
                    if (c < m) rows[(size_t)tt * pw + n + c] = (int8_t)(valid - 2 * __popcll(a[j] ^ b[j]));
                }
            }
#pragma unroll
            for (int j = 0; j < kEdgeBatch; ++j) cur[j] = nxt[j];
        }
    }
    __syncthreads();
    // the rows are consecutive in part (stride pw, a multiple of 16)
    const uint4* src = reinterpret_cast<const uint4*>(rows);
    uint4* dst = reinterpret_cast<uint4*>(part + blk.part_off);
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4* s4 = reinterpret_cast<const u32x4*>(src);
    u32x4* d4 = reinterpret_cast<u32x4*>(dst);
    for (int q = tid; q < ntiles * (pw >> 4); q += kMomentThreads) __builtin_nontemporal_store(s4[q], &d4[q]);
}

// grad_h[i] = scale * sum_s part[s][i] / R, grad_J[c] likewise.  A block
// covers kReduceSpan entries of one model, four per lane (one dword of each
// tile row); its four waves add every fourth tile row (16 loads in flight per
// lane) and wave 0 adds the four integer sums.
__global__ __launch_bounds__(256) void moments_reduce_kernel(const MomentSpan* __restrict__ spans,
                                                             const int8_t* __restrict__ part,
                                                             const float* __restrict__ scale,
                                                             float* __restrict__ grad_h, float* __restrict__ grad_J,
                                                             int64_t* __restrict__ sums) {
    static_assert(kReduceSpan == 4 * 64, "a wave covers the span, four entries per lane");
    __shared__ int4 red[4][64];
    const MomentSpan sp = spans[blockIdx.x];
    const MomentModel& md = sp.md;
    const int w = md.n + md.m, pw = md.pw;
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int e = sp.e0 + 4 * lane;  // this lane's first entry; rows are pw bytes, a multiple of 16
    int t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    auto add = [&](uint32_t d) {
        t0 += (int)__builtin_amdgcn_sbfe(d, 0, 8);
        t1 += (int)__builtin_amdgcn_sbfe(d, 8, 8);
        t2 += (int)__builtin_amdgcn_sbfe(d, 16, 8);
        t3 += (int)__builtin_amdgcn_sbfe(d, 24, 8);
    };
    if (e < w) {
        const uint32_t* p = reinterpret_cast<const uint32_t*>(part + md.part_off + e);
        const int64_t st = pw >> 2;  // row stride in dwords
        int s = g;
        for (; s + 60 < md.splits; s += 64) {
            uint32_t v[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) v[k] = p[(int64_t)(s + 4 * k) * st];
#pragma unroll
            for (int k = 0; k < 16; ++k) add(v[k]);
        }
        for (; s < md.splits; s += 4) add(p[(int64_t)s * st]);
    }
    red[g][lane] = make_int4(t0, t1, t2, t3);
    __syncthreads();
    if (g == 0 && e < w) {
        const float gs = scale ? scale[sp.model] : 1.0f;
        const float invR = md.R > 0 ? 1.0f / (float)md.R : 0.0f;
        const int4 r0 = red[0][lane], r1 = red[1][lane], r2 = red[2][lane], r3 = red[3][lane];
        const int64_t tot[4] = {(int64_t)r0.x + r1.x + r2.x + r3.x, (int64_t)r0.y + r1.y + r2.y + r3.y,
                                (int64_t)r0.z + r1.z + r2.z + r3.z, (int64_t)r0.w + r1.w + r2.w + r3.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ee = e + q;
            if (ee < w) {
                if (sums) sums[md.sum_off + ee] = tot[q];
                const float mean = (float)tot[q] * invR;
                if (ee < md.n) {
                    if (grad_h) grad_h[md.h_off + ee] = gs * mean;
                } else if (grad_J) {
                    grad_J[md.edge_off + (ee - md.n)] = gs * mean;
                }
            }
        }
    }
}

}  // namespace

hipError_t launch_moments_partial(const MomentBlock* d_blocks, int nblocks, int lds_bytes, const int8_t* samples,
                                  const int32_t* edges, int8_t* part, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    if (lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(moments_partial_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(moments_partial_kernel, dim3(nblocks), dim3(kMomentThreads), lds_bytes, stream, d_blocks,
                       samples, edges, part);
    return hipGetLastError();
}

hipError_t launch_moments_reduce(const MomentSpan* d_spans, int nspans, const int8_t* part, const float* scale,
                                 float* grad_h, float* grad_J, int64_t* sums, hipStream_t stream) {
    if (nspans <= 0) return hipSuccess;
    hipLaunchKernelGGL(moments_reduce_kernel, dim3(nspans), dim3(256), 0, stream, d_spans, part, scale, grad_h,
                       grad_J, sums);
    return hipGetLastError();
}

}  // namespace misa
