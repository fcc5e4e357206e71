// Segment kernel: float64 simulated annealing in sequential order, a run of
// variables per step of the read's lanes (gfx950).
//
// Same algorithm, float64 operation order and RNG stream as the reference
// core (dwave/samplers/sa/src/cpu_sa.cpp:93-197, xorshift128+ :27-34), and the
// same field form as sa_real_kernel.hip (f_v = h_v + sum J s, dE_v = -2 s_v f_v,
// a flip of v adds -2 J s_v to each neighbour's field).
//
// The host cuts 0..n-1 into runs of consecutive variables with no coupler
// between any two of them (a variable whose row repeats a neighbour or holds
// itself is a run of its own).  Inside a run no decision can change another
// member's field, so the reference's serial loop over the run differs from an
// all-at-once evaluation only through the random stream: a member draws a
// number iff its own dE lies strictly between 0 and the skip threshold
// (Metropolis; Gibbs: below the threshold), which depends on nothing but its
// field, so member k's draw is output number (count of drawing members before
// k) of the stream.  The L lanes of a read therefore hold the next L outputs of
// the read's xorshift128+ stream (lane i: output c + i, c the outputs consumed
// so far), and a run of K <= seg_kmax members is decided in one pass: lane k
// loads field and spin of member k, forms dE and the reference's double
// comparison with the output at its prefix count (a ds_bpermute inside the
// group), and the flips are applied member by member in increasing order
// (return-less float64 LDS atomics on the neighbours' fields, one member per
// instruction, so two updates of one field keep the reference's order).  The
// window then moves on by the number of draws, the new outputs generated in
// order at its end.  Loads of the next run (fields, spins, this lane's CSR
// entry of every member row) are issued after the current run's atomics,
// which LDS applies first.  Run starts and row bounds are wave-uniform and come
// through the scalar cache; with one read per wave (L = 64) its random stream
// and decisions are wave-uniform too.  Every row has at most L entries.
//
// LDS regions (byte offsets in SaArgs):
//   0                 fields, read-major: F[slot * fstride + v] (float64)
//   lds_off_J_spins   spins, read-major: S[slot * sstride + v] (int8)
//   lds_off_col/_J    CSR columns (u16) and weights (f64) + 64 spare, when staged
#include "sa_real_block.h"
#include "xorshift.h"

namespace misa {

namespace seg_detail {

// Bit 0 of every group of L lanes.
template <int L>
__device__ constexpr uint64_t group_bits() {
    uint64_t b = 0;
    for (int g = 0; g < 64; g += L) b |= 1ull << g;
    return b;
}

template <int L>
__device__ constexpr uint64_t lane_mask() {
    return L == 64 ? ~0ull : ((1ull << L) - 1);
}

// xorshift128+ (cpu_sa.cpp:27-34) without committing the step.
struct Step {
    uint64_t s0, s1, out;
};
__device__ __forceinline__ Step xs_step(uint64_t s0, uint64_t s1) {
    uint64_t a = s0;
    const uint64_t b = s1;
    a ^= a << 23;
    const uint64_t n = a ^ b ^ (a >> 17) ^ (b >> 26);
    return {b, n, n + b};
}

// Wave-uniform load through the scalar cache (the address must be uniform).
template <typename T>
__device__ __forceinline__ T sload(const T* p) {
    return *reinterpret_cast<const __attribute__((address_space(4))) T*>(reinterpret_cast<uintptr_t>(p));
}

constexpr int SERIAL_RUN = (int)0x80000000u;

}  // namespace seg_detail

// One block's share of a segment launch: reads blk*rpb ... blk*rpb + rpb - 1.
// TPL > 0 (one read per wave only): the members' rows come from the run
// tiles (rows padded to 1 << TPL entries, 64 >> TPL members per 64-entry
// slot), one coalesced load per slot, and the flips are applied slot by slot
// over the set bits of the flip mask.
// JMP (with the tiles): lane l runs its own sub-stream, the outputs l, l + 64,
// l + 128, ... of the read's stream (the first reached by l steps, each next
// by a 64-step jump through the table staged at lds_off_jump).  The lanes
// hold a circular window: lane l holds the one output o_l = l (mod 64) of
// [c, w) (c: outputs consumed, w - c >= 32, at most the 32 draws of a run),
// and when fewer than 32 are left every lane holding a consumed output jumps
// to the output 64 on, so w = c + 64 again.  A run's draw after p others is
// the output held by lane (c + p) mod 64.
template <int PROP, bool CSR_LDS, int LG, bool COUNT, int TPL = 0, bool JMP = false>
__device__ __forceinline__ void seg_block(const SaArgs& a, const int blk, unsigned char* smem) {
    using namespace seg_detail;
    constexpr int L = 1 << LG;
    constexpr int KMAX = seg_kmax(LG);
    constexpr bool TILES = TPL > 0;
    static_assert(!TILES || (LG == 6 && TPL >= 4 && TPL <= 6), "tiles: one read per wave, rows of 16..64");
    static_assert(!JMP || TILES, "per-lane streams come with the tiles");
    static_assert(!JMP || KMAX <= 32, "a run draws at most 32 outputs");
    constexpr int TG = TILES ? 64 >> TPL : 1;          // members per slot
    constexpr int TQ = TILES ? (KMAX + TG - 1) / TG : 1;  // slots of a run
    constexpr uint64_t GB = group_bits<L>();
    constexpr uint64_t LM = lane_mask<L>();
    constexpr bool WAVE = L == 64;  // one read per wave
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & (L - 1);
    const int gbase = lane & ~(L - 1);
    const int slot = wave * (64 >> LG) + (lane >> LG);
    const int n = a.n;
    const bool own = slot < a.rpb;
    const int rlocal = blk * a.rpb + slot;
    const bool act = own && rlocal < a.R;
    const int r = a.r0 + rlocal;

    double* Fr = reinterpret_cast<double*>(smem) + (size_t)slot * a.fstride;
    int8_t* Sr = reinterpret_cast<int8_t*>(smem + a.lds_off_J_spins) + (size_t)slot * a.sstride;
    const uint16_t* ccol = reinterpret_cast<const uint16_t*>(smem + a.lds_off_col);
    const double* cJ = reinterpret_cast<const double*>(smem + a.lds_off_J);
    auto col_at = [&](const int k) -> int { return CSR_LDS ? (int)ccol[k] : a.col[k]; };
    auto w_at = [&](const int k) -> double { return CSR_LDS ? cJ[k] : a.Jd[k]; };

    // ---- stage the adjacency, spins in, then fields (cpu_sa.cpp:62-68) ---
    if constexpr (CSR_LDS) {
        uint16_t* wcol = reinterpret_cast<uint16_t*>(smem + a.lds_off_col);
        double* wJ = reinterpret_cast<double*>(smem + a.lds_off_J);
        for (int64_t k = threadIdx.x; k < a.m2 + 64; k += blockDim.x) {
            wcol[k] = k < a.m2 ? (uint16_t)a.col[k] : (uint16_t)0;
            wJ[k] = k < a.m2 ? a.Jd[k] : 0.0;
        }
    }
    if (own)
        for (int v = j; v < n; v += L) Sr[v] = act ? a.states[(size_t)r * n + v] : (int8_t)1;
    const uint4* Jt = reinterpret_cast<const uint4*>(smem + a.lds_off_jump);
    if constexpr (JMP)
        for (int t = threadIdx.x; t < kJumpEntries; t += blockDim.x)
            reinterpret_cast<uint4*>(smem + a.lds_off_jump)[t] = reinterpret_cast<const uint4*>(a.chain_jump)[t];
    __syncthreads();
    if (own)
        for (int v = j; v < n; v += L) {
            const int kb = a.rowptr[v], ke = a.rowptr[v + 1];
            double e = a.h[v];
            for (int k = kb; k < ke; ++k) e += (double)Sr[col_at(k)] * w_at(k);
            Fr[v] = e;
        }
    __syncthreads();
    if (!act) return;

    // ---- the read's stream: lane j holds output c + j ----------------------
    uint64_t s0 = a.seeds[r];
    if constexpr (WAVE) s0 = ((uint64_t)__builtin_amdgcn_readfirstlane((int)(s0 >> 32)) << 32) |
                             (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)s0);
    if (s0 == 0) s0 = ~0ull;  // cpu_sa.cpp:283-284
    uint64_t s1 = 0;
    uint32_t wlo = 0, whi = 0;
    // per-lane streams: lane j's state before its output o (o = j first)
    uint32_t q0 = (uint32_t)s0, q1 = (uint32_t)(s0 >> 32), q2 = 0, q3 = 0;
    uint32_t o = (uint32_t)j, cpos = 0, wend = 64;
    if constexpr (JMP) {
        for (int i = 0; i < 63; ++i) {
            const XsStep st = xs_peek(q0, q1, q2, q3);
            const bool go = i < j;
            q0 = go ? q2 : q0;
            q1 = go ? q3 : q1;
            q2 = go ? st.nlo : q2;
            q3 = go ? st.nhi : q3;
        }
        const XsStep st = xs_peek(q0, q1, q2, q3);
        wlo = st.olo;
        whi = st.ohi;
    } else {
        for (int i = 0; i < L; ++i) {
            const Step st = xs_step(s0, s1);
            s0 = st.s0;
            s1 = st.s1;
            if (i == j) {
                wlo = (uint32_t)st.out;
                whi = (uint32_t)(st.out >> 32);
            }
        }
    }
    // (s0, s1) now produce output c + L

    unsigned long long nflip = 0, ndeg = 0;
    const int nseg = a.nseg;
    const uint64_t below = (1ull << j) - 1;

    // registers of the run being decided
    int v0 = 0, K = 0;
    bool serial = false;
    double f = 0.0;
    int s = 1;
    int rs[KMAX + 1];  // row starts of members 0..KMAX (wave-uniform)
    int cu[TILES ? 1 : KMAX];  // this lane's CSR entry of every member row
    double cw[TILES ? 1 : KMAX];
    int tc[TQ];        // tiles: this lane's entry of every slot
    double tj[TQ];
    // tiles: a run's bounds (start with the serial flag, end, tile offset,
    // row bounds of its first member) are read by vector loads during the
    // run before and made uniform when the run starts (a scalar load would
    // hold up the run's first LDS loads)
    int raw[5];
    auto read_bounds = [&](const int sg) {
        int i0 = sg, i1 = 4 * sg;
        asm("" : "+v"(i0), "+v"(i1));
        raw[0] = a.seg[i0];
        raw[1] = a.seg[i0 + 1];
        raw[2] = a.tile_off[i1];
        raw[3] = a.tile_off[i1 + 1];
        raw[4] = a.tile_off[i1 + 2];
    };

    auto load_run = [&](const int sg) {
        int w0, w1;
        if constexpr (TILES) {
            w0 = __builtin_amdgcn_readfirstlane(raw[0]);
            w1 = __builtin_amdgcn_readfirstlane(raw[1]);
        } else {
            w0 = sload(a.seg + sg);
            w1 = sload(a.seg + sg + 1);
        }
        v0 = w0 & ~SERIAL_RUN;
        K = (w1 & ~SERIAL_RUN) - v0;
        serial = w0 < 0;
        const int vj = v0 + (j < K ? j : 0);
        f = Fr[vj];
        s = Sr[vj];
        if constexpr (TILES) {
            // the serial path reads its row from the CSR
            const int toff = __builtin_amdgcn_readfirstlane(raw[2]);
            rs[0] = __builtin_amdgcn_readfirstlane(raw[3]);
            rs[1] = __builtin_amdgcn_readfirstlane(raw[4]);
            #pragma unroll
            for (int q = 0; q < TQ; ++q) {
                if (q * TG < K) {
                    tc[q] = a.tile_col[toff + q * 64 + j];
                    tj[q] = a.tile_J[toff + q * 64 + j];
                }
            }
            read_bounds(sg + 1 == nseg ? 0 : sg + 1);
        } else {
            #pragma unroll
            for (int t = 0; t <= KMAX; ++t) rs[t] = sload(a.rowptr_pad + v0 + t);  // padded past n
            #pragma unroll
            for (int t = 0; t < KMAX; ++t) {
                if (t < K) {
                    cu[t] = col_at(rs[t] + j);  // past the row's end: the next rows or spare entries
                    cw[t] = w_at(rs[t] + j);
                }
            }
        }
    };

    if constexpr (TILES) read_bounds(0);
    load_run(0);
    int sg = 0;
    for (int bi = 0; bi < a.nbetas; ++bi) {
        const double beta = a.betas[bi];
        const double thr = a.thr[bi];  // 44.36142 / beta, cpu_sa.cpp:131
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            for (int c = 0; c < nseg; ++c) {
                // ---- decide every member (cpu_sa.cpp:140-166) ----------------
                const double de = (double)(-2 * s) * f;
                const bool go = j < K && !(de >= thr);
                bool draw, free_flip;
                if constexpr (PROP == 1) {
                    free_flip = go && (de <= 0.0);
                    draw = go && !(de <= 0.0);
                } else {
                    free_flip = false;
                    draw = go;
                }
                const uint64_t gdraw = (__builtin_amdgcn_ballot_w64(draw) >> gbase) & LM;
                const int d = __popcll(gdraw);
                const int src = JMP ? (int)((cpos + (uint32_t)__popcll(gdraw & below)) & 63u) << 2
                                    : (gbase + __popcll(gdraw & below)) << 2;
                const uint64_t x = (uint64_t)(uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)whi) << 32 |
                                   (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)wlo);
                bool flip;
                if constexpr (PROP == 1)
                    flip = free_flip || (draw && real_detail::metropolis_accept(-de * beta, x));  // :154
                else
                    flip = draw && real_detail::gibbs_accept(de * beta, x);  // :163
                const uint64_t flipm = __builtin_amdgcn_ballot_w64(flip);
                const uint64_t negm = __builtin_amdgcn_ballot_w64(s < 0);
                if (flip) Sr[v0 + j] = (int8_t)(-s);
                // ---- neighbour updates, members in increasing order ----------
                if (serial) {
                    // one variable whose row repeats a neighbour or holds itself:
                    // one lane, CSR order (the reference's float order)
                    if (j == 0 && flip) {
                        const double m2s = (double)(-2 * s);
                        for (int k = rs[0]; k < rs[1]; ++k) Fr[col_at(k)] += m2s * w_at(k);
                        if constexpr (COUNT) {
                            nflip += 1;
                            ndeg += (unsigned)(rs[1] - rs[0]);
                        }
                    }
                } else if constexpr (TILES) {
                    // slot by slot, the slot's flipped members in increasing order
                    #pragma unroll
                    for (int q = 0; q < TQ; ++q) {
                        uint32_t bits = (uint32_t)(flipm >> (q * TG)) & ((1u << TG) - 1u);
                        while (bits) {
                            const int gi = __builtin_ctz(bits);
                            bits &= bits - 1u;
                            const int t = q * TG + gi;
                            const double m2s = ((negm >> t) & 1) ? 2.0 : -2.0;  // -2 s_t
                            const bool mine = (j >> TPL) == gi && tc[q] >= 0;
                            if (mine)
                                __hip_atomic_fetch_add(&Fr[tc[q]], m2s * tj[q], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            if constexpr (COUNT) {
                                nflip += 1;
                                ndeg += (unsigned)__popcll(__builtin_amdgcn_ballot_w64(mine));
                            }
                        }
                    }
                } else {
                    const uint32_t gflip = (uint32_t)(flipm >> gbase);
                    const uint32_t gneg = (uint32_t)(negm >> gbase);
                    #pragma unroll
                    for (int t = 0; t < KMAX; ++t) {
                        if (flipm & (GB << t)) {
                            const int deg = rs[t + 1] - rs[t];
                            if constexpr (WAVE) {
                                const double m2s = ((negm >> t) & 1) ? 2.0 : -2.0;  // -2 s_t
                                if (j < deg)
                                    __hip_atomic_fetch_add(&Fr[cu[t]], m2s * cw[t], __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                                if constexpr (COUNT) {
                                    nflip += 1;
                                    ndeg += (unsigned)deg;
                                }
                            } else {
                                const bool mine = (gflip >> t) & 1;
                                if (mine && j < deg) {
                                    const double m2s = ((gneg >> t) & 1) ? 2.0 : -2.0;  // -2 s_t
                                    __hip_atomic_fetch_add(&Fr[cu[t]], m2s * cw[t], __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                                }
                                if constexpr (COUNT) {
                                    nflip += mine ? 1u : 0u;
                                    ndeg += mine ? (unsigned)deg : 0u;
                                }
                            }
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
                // ---- next run's loads, issued behind this run's updates ------
                sg = sg + 1 == nseg ? 0 : sg + 1;
                load_run(sg);
                // ---- move the window on by d outputs ------------------------
                if constexpr (JMP) {
                    cpos += (uint32_t)d;
                    if (wend - cpos < 32u) {
                        // every lane holding a consumed output moves 64 on
                        const bool stale = (int32_t)(o - cpos) < 0;
                        uint32_t x0 = q0, x1 = q1, y0 = q2, y1 = q3;
                        xs_jump(x0, x1, y0, y1, Jt);
                        const XsStep st = xs_peek(x0, x1, y0, y1);
                        q0 = stale ? x0 : q0;
                        q1 = stale ? x1 : q1;
                        q2 = stale ? y0 : q2;
                        q3 = stale ? y1 : q3;
                        wlo = stale ? st.olo : wlo;
                        whi = stale ? st.ohi : whi;
                        o = stale ? o + 64u : o;
                        wend = cpos + 64u;
                    }
                    continue;
                }
                const int from = (gbase + ((j + d) & (L - 1))) << 2;
                uint32_t nlo = (uint32_t)__builtin_amdgcn_ds_bpermute(from, (int)wlo);
                uint32_t nhi = (uint32_t)__builtin_amdgcn_ds_bpermute(from, (int)whi);
                if constexpr (WAVE) {
                    for (int t = 0; t < d; ++t) {
                        const Step st = xs_step(s0, s1);
                        s0 = st.s0;
                        s1 = st.s1;
                        if (j == L - d + t) {
                            nlo = (uint32_t)st.out;
                            nhi = (uint32_t)(st.out >> 32);
                        }
                    }
                } else {
                    for (int t = 0; __builtin_amdgcn_ballot_w64(t < d) != 0; ++t) {
                        const Step st = xs_step(s0, s1);
                        const bool take = t < d;
                        s0 = take ? st.s0 : s0;
                        s1 = take ? st.s1 : s1;
                        const bool put = j == L - d + t;
                        nlo = put ? (uint32_t)st.out : nlo;
                        nhi = put ? (uint32_t)(st.out >> 32) : nhi;
                    }
                }
                wlo = nlo;
                whi = nhi;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- spins out and energy (cpu_sa.cpp:210-227) -------------------------
    for (int v = j; v < n; v += L) a.states[(size_t)r * n + v] = Sr[v];
    if (j == 0) {
        double en = 0.0;
        for (int v = 0; v < n; ++v) en += (double)Sr[v] * a.h[v];
        for (int64_t c = 0; c < a.m; ++c) en += (double)Sr[a.cs[c]] * a.cw[c] * (double)Sr[a.ce[c]];
        a.energies[r] = en;
        if (COUNT && a.counters) {
            a.counters[2 * (size_t)r] = nflip;
            a.counters[2 * (size_t)r + 1] = ndeg;
        }
    }
}

// ---------------------------------------------------------------------------
// Plain launch
// ---------------------------------------------------------------------------

template <int PROP, bool CSR_LDS, int LG, bool COUNT, int TPL, bool JMP>
__global__ __launch_bounds__(512) void sa_seg_kernel(SaArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    seg_block<PROP, CSR_LDS, LG, COUNT, TPL, JMP>(a, blockIdx.x, smem);
}

template <int PROP, bool CSR, int LG, bool COUNT, int TPL, bool JMP = false>
static hipError_t launch_seg_tiles(const SaArgs& a, int nblocks, hipStream_t stream) {
    auto fn = sa_seg_kernel<PROP, CSR, LG, COUNT, TPL, JMP>;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(64 * a.waves_per_block), a.lds_bytes, stream, a);
    return hipGetLastError();
}

template <int PROP, bool CSR, int LG, bool COUNT>
static hipError_t launch_seg_variant(const SaArgs& a, int nblocks, hipStream_t stream) {
    if constexpr (LG == 6) {
        if (a.tile_col && a.chain_jump) {
            if (a.tile_plog2 == 4) return launch_seg_tiles<PROP, CSR, LG, COUNT, 4, true>(a, nblocks, stream);
            if (a.tile_plog2 == 5) return launch_seg_tiles<PROP, CSR, LG, COUNT, 5, true>(a, nblocks, stream);
            if (a.tile_plog2 == 6) return launch_seg_tiles<PROP, CSR, LG, COUNT, 6, true>(a, nblocks, stream);
            return hipErrorInvalidValue;
        }
        if (a.tile_col) {
            if (a.tile_plog2 == 4) return launch_seg_tiles<PROP, CSR, LG, COUNT, 4>(a, nblocks, stream);
            if (a.tile_plog2 == 5) return launch_seg_tiles<PROP, CSR, LG, COUNT, 5>(a, nblocks, stream);
            if (a.tile_plog2 == 6) return launch_seg_tiles<PROP, CSR, LG, COUNT, 6>(a, nblocks, stream);
            return hipErrorInvalidValue;
        }
    }
    return launch_seg_tiles<PROP, CSR, LG, COUNT, 0>(a, nblocks, stream);
}

template <int PROP, bool CSR, bool COUNT>
static hipError_t launch_seg_lanes(const SaArgs& a, int nb, hipStream_t s) {
    switch (a.lanes_log2) {
        case 3: return launch_seg_variant<PROP, CSR, 3, COUNT>(a, nb, s);
        case 4: return launch_seg_variant<PROP, CSR, 4, COUNT>(a, nb, s);
        case 5: return launch_seg_variant<PROP, CSR, 5, COUNT>(a, nb, s);
        case 6: return launch_seg_variant<PROP, CSR, 6, COUNT>(a, nb, s);
        default: return hipErrorInvalidValue;
    }
}

template <int PROP>
static hipError_t launch_seg_prop(const SaArgs& a, bool csr, int nb, hipStream_t s) {
    const bool count = a.counters != nullptr;
    if (csr) return count ? launch_seg_lanes<PROP, true, true>(a, nb, s) : launch_seg_lanes<PROP, true, false>(a, nb, s);
    return count ? launch_seg_lanes<PROP, false, true>(a, nb, s) : launch_seg_lanes<PROP, false, false>(a, nb, s);
}

hipError_t launch_sa_seg(const SaArgs& a, int proposal, bool csr_in_lds, int nblocks, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    return proposal == 1 ? launch_seg_prop<1>(a, csr_in_lds, nblocks, stream)
                         : launch_seg_prop<0>(a, csr_in_lds, nblocks, stream);
}

// ---------------------------------------------------------------------------
// Batched launch: a persistent grid over a device work queue
// ---------------------------------------------------------------------------

// One read per wave: with or without the run tiles (rows of 16, 32 or 64),
// with the tiles with or without per-lane streams.
template <int PROP, bool CSR>
__device__ __forceinline__ void seg_wave(const SaArgs& a, const int blk, unsigned char* smem, bool cnt) {
    if (a.tile_col && a.chain_jump && a.tile_plog2 == 4)
        cnt ? seg_block<PROP, CSR, 6, true, 4, true>(a, blk, smem) : seg_block<PROP, CSR, 6, false, 4, true>(a, blk, smem);
    else if (a.tile_col && a.chain_jump && a.tile_plog2 == 5)
        cnt ? seg_block<PROP, CSR, 6, true, 5, true>(a, blk, smem) : seg_block<PROP, CSR, 6, false, 5, true>(a, blk, smem);
    else if (a.tile_col && a.chain_jump && a.tile_plog2 == 6)
        cnt ? seg_block<PROP, CSR, 6, true, 6, true>(a, blk, smem) : seg_block<PROP, CSR, 6, false, 6, true>(a, blk, smem);
    else if (a.tile_col && a.tile_plog2 == 4)
        cnt ? seg_block<PROP, CSR, 6, true, 4>(a, blk, smem) : seg_block<PROP, CSR, 6, false, 4>(a, blk, smem);
    else if (a.tile_col && a.tile_plog2 == 5)
        cnt ? seg_block<PROP, CSR, 6, true, 5>(a, blk, smem) : seg_block<PROP, CSR, 6, false, 5>(a, blk, smem);
    else if (a.tile_col && a.tile_plog2 == 6)
        cnt ? seg_block<PROP, CSR, 6, true, 6>(a, blk, smem) : seg_block<PROP, CSR, 6, false, 6>(a, blk, smem);
    else
        cnt ? seg_block<PROP, CSR, 6, true>(a, blk, smem) : seg_block<PROP, CSR, 6, false>(a, blk, smem);
}

// Every block takes items in queue order (the host sorts them longest chain
// first) until the queue is empty; the last block to leave resets both
// counters, so the next launch (or graph replay) starts from item 0.  Items
// whose counters pointer is null run without counting.
template <int ORDER, int PROP>
__global__ __launch_bounds__(kBatchThreads) void sa_batch_kernel(const BatchItem* __restrict__ items, int nitems,
                                                                 unsigned* queue) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int next_item;
    for (;;) {
        if (threadIdx.x == 0)
            next_item = (int)__hip_atomic_fetch_add(&queue[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int i = next_item;
        __syncthreads();
        if (i >= nitems) break;
        const BatchItem& it = items[i];
        if constexpr (ORDER == 0) {
            if (it.kind == 1) {
                const int lg = it.args.lanes_log2;
                const bool cnt = it.args.counters != nullptr;
                if (it.csr_in_lds) {
                    if (lg == 4) cnt ? seg_block<PROP, true, 4, true>(it.args, it.block, smem)
                                     : seg_block<PROP, true, 4, false>(it.args, it.block, smem);
                    else if (lg == 5) cnt ? seg_block<PROP, true, 5, true>(it.args, it.block, smem)
                                          : seg_block<PROP, true, 5, false>(it.args, it.block, smem);
                    else if (lg == 6) seg_wave<PROP, true>(it.args, it.block, smem, cnt);
                    else cnt ? seg_block<PROP, true, 3, true>(it.args, it.block, smem)
                             : seg_block<PROP, true, 3, false>(it.args, it.block, smem);
                } else {
                    if (lg == 4) cnt ? seg_block<PROP, false, 4, true>(it.args, it.block, smem)
                                     : seg_block<PROP, false, 4, false>(it.args, it.block, smem);
                    else if (lg == 5) cnt ? seg_block<PROP, false, 5, true>(it.args, it.block, smem)
                                          : seg_block<PROP, false, 5, false>(it.args, it.block, smem);
                    else if (lg == 6) seg_wave<PROP, false>(it.args, it.block, smem, cnt);
                    else cnt ? seg_block<PROP, false, 3, true>(it.args, it.block, smem)
                             : seg_block<PROP, false, 3, false>(it.args, it.block, smem);
                }
            } else if (it.csr_in_lds) {
                real_block<ORDER, PROP, true>(it.args, it.block, smem);
            } else {
                real_block<ORDER, PROP, false>(it.args, it.block, smem);
            }
        } else {
            if (it.csr_in_lds) real_block<ORDER, PROP, true>(it.args, it.block, smem);
            else real_block<ORDER, PROP, false>(it.args, it.block, smem);
        }
        // every wave is done with this item's LDS before the next is staged
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const unsigned done = __hip_atomic_fetch_add(&queue[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (done + 1 == gridDim.x) {
            // every block has left the queue loop: reset for the next launch
            __hip_atomic_store(&queue[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&queue[1], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int ORDER, int PROP>
static hipError_t launch_batch_variant(const BatchItem* d_items, int nitems, unsigned* queue, int grid, unsigned lds,
                                       hipStream_t stream) {
    auto fn = sa_batch_kernel<ORDER, PROP>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(grid), dim3(kBatchThreads), lds, stream, d_items, nitems, queue);
    return hipGetLastError();
}

template <int ORDER, int PROP>
static hipError_t batch_occupancy(unsigned lds, int* out) {
    auto fn = sa_batch_kernel<ORDER, PROP>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, reinterpret_cast<const void*>(fn), kBatchThreads, lds);
}

hipError_t batch_blocks_per_cu(int order, int proposal, unsigned lds_bytes, int* out) {
    *out = 1;
    if (order == 0)
        return proposal == 1 ? batch_occupancy<0, 1>(lds_bytes, out) : batch_occupancy<0, 0>(lds_bytes, out);
    return proposal == 1 ? batch_occupancy<1, 1>(lds_bytes, out) : batch_occupancy<1, 0>(lds_bytes, out);
}

hipError_t launch_sa_batch(const BatchItem* d_items, int nitems, unsigned* queue, int grid, unsigned lds_bytes,
                           int order, int proposal, hipStream_t stream) {
    if (nitems <= 0 || grid <= 0) return hipSuccess;
    if (order == 0)
        return proposal == 1 ? launch_batch_variant<0, 1>(d_items, nitems, queue, grid, lds_bytes, stream)
                             : launch_batch_variant<0, 0>(d_items, nitems, queue, grid, lds_bytes, stream);
    return proposal == 1 ? launch_batch_variant<1, 1>(d_items, nitems, queue, grid, lds_bytes, stream)
                         : launch_batch_variant<1, 0>(d_items, nitems, queue, grid, lds_bytes, stream);
}

}  // namespace misa
