// Integer chain-run sweep kernel for gfx950 (int8 words, sequential order).
//
// Same algorithm, RNG stream and results as the reference core
// (dwave/samplers/sa/src/cpu_sa.cpp:93-197, xorshift128+ :27-34) and as
// sa_int_kernel; what differs is how one read's sequential chain is
// evaluated.  The host cuts the variables 0..n-1 into runs of at most
// kChainK consecutive variables in which a member is coupled to no earlier
// member except its immediate predecessor.  A member's word at its turn is
// then its word at the start of the run, plus -4 J(v-1, v) s_{v-1} when its
// predecessor flipped.  Sixteen lanes (one DPP row) hold a read and lane k
// takes member k of each run:
//  1. each lane loads its member's word and the decision-table entries for
//     both outcomes of its predecessor (T0: kept, T1: flipped);
//  2. each lane compares both entries with the read's next kChainK
//     xorshift128+ outputs x_1..x_K (the same in each lane of the group):
//     bit K-1-i of m_h is x_{i+1} < T_h, all ones when T_h = ~0 (flip
//     without a draw); T_h = 0 (reject without a draw) compares false
//     everywhere.  The kernel's table holds per word -T, the high word of T
//     and those class bits.  The outputs come from a per-read ring in LDS that the
//     wave refills 256 at a time (ring variant), or every lane of the group
//     steps the generator kChainK times and keeps the state words in LDS.
//     Few-threshold form (a beta's drawing words that a variable can hold
//     have at most kChainFewT distinct T, as on a +-1 lattice): lane k of a
//     group holds x_{15-k} only, and one ballot per threshold T_a gives every
//     group's 15 bits x_{i+1} < T_a at once; a lane packs its group's bits of
//     the thresholds into 16-bit segments of one 64-bit word and takes an
//     entry's bits with a shift by 16 a (the table holds 16 a and the class
//     bits);
//  3. a serial pass over the members, each member's bits broadcast along the
//     row (DPP row_newbcast): member k flips iff the bit of m_pf for d draws
//     so far is set, pf the flip of member k-1; d grows by one when the
//     member's entry asks for a draw;
//  4. the read's position advances by d draws, and each flipped member adds
//     -4 J s to its neighbours' words and toggles its own spin bit with
//     return-less LDS atomics (the adds commute).
// Words are BIAS + 2 f + b as in sa_int_kernel (b = 1 for spin -1).  Four
// groups per wave, four waves per block (16 reads); word (v, read) sits at
// byte v * 16 + 4 g + w (g the group, w the wave), so the four reads of a
// wave use four different dwords of a row.
#include <type_traits>

#include "sa_engine.h"
#include "xorshift.h"

namespace misa {

namespace {

constexpr int kBias = 128;

__device__ __forceinline__ void lds_add(unsigned char* smem, uint32_t byte, uint32_t val) {
    uint32_t* w = static_cast<uint32_t*>(__builtin_assume_aligned(smem + byte, 4));
    __hip_atomic_fetch_add(w, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The same add at LDS byte address `byte`: the kernel declares no static LDS
// (the launch checks), so its dynamic LDS starts at address 0 and the offset
// is the address, with no base to add.
__device__ __forceinline__ void lds_add_at(uint32_t byte, uint32_t val) {
    auto* w = reinterpret_cast<__attribute__((address_space(3))) uint32_t*>((size_t)byte);
    __hip_atomic_fetch_add(w, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Serial pass over the members KK..kChainK-1.  Db (per lane): bits 0..14
// m0, bit 15 "entry T0 draws", bits 16..30 m1, bit 31 "entry T1 draws"; zero
// for lanes past the run; bit 14 - i of m_h is the outcome of the draw with
// i draws before it.  e = 14 - (draws so far); pf: 16 when the previous
// member flipped; flips: bit KK set when member KK flips.
template <int KK>
__device__ __forceinline__ void resolve(uint32_t Db, uint32_t& e, uint32_t& pf, uint32_t& flips) {
    const uint32_t Dk = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)Db, 0x150 + KK, 0xf, 0xf, true);
    const uint32_t Ds = Dk >> pf;
    const uint32_t fl = (Ds >> e) & 1u;
    e -= (Ds >> 15) & 1u;
    pf = fl << 4;
    flips |= fl << KK;
    if constexpr (KK + 1 < kChainK) resolve<KK + 1>(Db, e, pf, flips);
}

// (a & b) | c in one v_bitop3 (truth table indexed by S0 S1 S2 = a b c)
__device__ __forceinline__ uint32_t and_or(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xea);
}

// Appends the outcome of x < T (unsigned 64-bit) below the bits of m, from
// nT = -T and the high word of T: the borrow of x - T is bit 63 of
// (~x & T) | (~(x ^ T) & (x - T)), one v_bitop3 on the high words (truth
// table indexed by S0 S1 S2 = x T d); x + nT is one 64-bit add.
__device__ __forceinline__ uint32_t push_less(uint32_t m, uint64_t x, uint32_t Thi, uint64_t nT) {
    const uint64_t d = x + nT;
    const uint32_t b = __builtin_amdgcn_bitop3_b32((uint32_t)(x >> 32), Thi, (uint32_t)(d >> 32), 0x8e);
    return __builtin_amdgcn_alignbit(m, b, 31);
}

// A variable's record: dword 0 its info, dwords 1..NE its neighbour entries
// (loaded with no unused words, so none of them is a dead register that a
// later write would have to wait for).
template <int NE>
struct Rec {
    uint32_t w[NE + 1];
};
template <int NE>
__device__ __forceinline__ Rec<NE> load_rec(const uint32_t* __restrict__ rec, int v) {
    const uint32_t* p = rec + (size_t)v * kChainRecDwords;
    Rec<NE> r;
    const uint4 a = *reinterpret_cast<const uint4*>(p);
    r.w[0] = a.x;
    r.w[1] = a.y;
    r.w[2] = a.z;
    r.w[3] = a.w;
    if constexpr (NE == 4) {
        r.w[4] = p[4];
    } else if constexpr (NE == 6) {
        const uint2 b = *reinterpret_cast<const uint2*>(p + 4);
        r.w[4] = b.x;
        r.w[5] = b.y;
        r.w[6] = p[6];
    } else {
        static_assert(NE == 7, "records hold 4, 6 or 7 entries");
        const uint4 b = *reinterpret_cast<const uint4*>(p + 4);
        r.w[4] = b.x;
        r.w[5] = b.y;
        r.w[6] = b.z;
        r.w[7] = b.w;
    }
    return r;
}

}  // namespace

// SLOT (rows of at most 6 entries): records come from the run-major slot
// array, dword 7 naming the member's word row and whether the lane is past
// its run, so a run needs no bounds and a record no address arithmetic.
template <int NE, int NCRIT, bool COUNT, bool RING, bool FEW, bool SLOT>
__global__ __launch_bounds__(256) void sa_chain_kernel(SaArgs a, const uint32_t* __restrict__ rec,
                                                       const int* __restrict__ runs,
                                                       const uint4* __restrict__ jump) {
    extern __shared__ __align__(16) unsigned char smem[];
    static_assert(kChainK <= 15 && kChainDraw == 1u << 15, "two 15-bit masks and two draw bits per member fill one dword");
    static_assert(NCRIT <= NE, "critical entries are a prefix of the entries");
    static_assert(!FEW || RING, "the few-threshold form reads the output rings");
    static_assert(!SLOT || NE <= kChainRecDwords - 2, "slot records keep dword 7 for the member");
    constexpr int NR = SLOT ? kChainRecDwords - 1 : NE;  // record dwords loaded past dword 0
    static_assert(kChainFewT == 3, "three 16-bit segments and a zero one fill the packed bits");
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int k = lane & 15;  // member index within a run
    const int g = lane >> 4;  // group (read) within the wave
    const int slot = wave * 4 + g;
    const int n = a.n;
    const int rlocal = blockIdx.x * 16 + slot;
    const bool act = rlocal < a.R;
    // a slot past the launch's reads runs a copy of its last read and stores
    // nothing (the ring refills need every lane of the wave)
    const int r = a.r0 + (act ? rlocal : a.R - 1);
    const uint32_t pdw = (uint32_t)g * 4u;      // this read's dword in a row of 16 bytes
    const uint32_t psh = (uint32_t)wave * 8u;   // its byte in that dword
    const uint32_t pbyte = pdw + (uint32_t)wave;
    uint4* Tw = reinterpret_cast<uint4*>(smem + a.lds_off_T) + (size_t)wave * a.lut_width;
    uint2* Tw2 = reinterpret_cast<uint2*>(smem + a.lds_off_T) + (size_t)wave * a.lut_width;
    // few-threshold form: the output this lane compares (x_{15-k}; lane 15
    // reads x_16, whose bit is masked off) and its group's bit offset
    const int wk = k < kChainK ? kChainK - 1 - k : kChainK;
    const uint32_t gsh = 16u * (uint32_t)g;
    uint4* Jt = reinterpret_cast<uint4*>(smem + a.lds_off_jump);

    // ---- spins in, then fields h_v + sum_k J_k s_k (integers: order-free) --
    if constexpr (RING)
        for (int t = threadIdx.x; t < kJumpEntries; t += blockDim.x) Jt[t] = jump[t];
    for (int v = k; v < n; v += 16) {
        const int s = (int)a.states[(size_t)r * n + v];
        smem[v * 16 + pbyte] = (unsigned char)(kBias + (s < 0 ? 1 : 0));
    }
    __syncthreads();
    for (int v = k; v < n; v += 16) {
        const int kb = a.rowptr[v], ke = a.rowptr[v + 1];
        int e = (int)a.h[v];
        for (int kk = kb; kk < ke; ++kk) e += (1 - 2 * (smem[a.col[kk] * 16 + pbyte] & 1)) * (int)a.Ji[kk];
        const int b = smem[v * 16 + pbyte] & 1;  // bit 0 is unchanged by this write
        smem[v * 16 + pbyte] = (unsigned char)(kBias + 2 * e + b);
    }
    __syncthreads();
    if constexpr (!RING)
        if (!act) return;

    // RNG.  Without the ring: the read's state (s0, s1) in every lane of the
    // group.  With the ring: output t of read G of this wave (t = 1, 2, ...)
    // sits at ring_G[(t - 1) mod kRingSize]; lane L holds, for each of the
    // wave's four reads, a sub-stream positioned kRingM * L outputs into the
    // read's next block of kRingBatch outputs.  pm / cc: outputs produced /
    // drawn (the same in every lane of a group).
    uint32_t alo = 0, ahi = 0, blo = 0, bhi = 0;
    uint32_t q0[4], q1[4], q2[4], q3[4];
    uint32_t pm = 0, cc = 0;
    uint64_t* Xs = reinterpret_cast<uint64_t*>(smem + a.lds_off_rng) + (size_t)slot * kChainRngStride;
    uint64_t* ringw = reinterpret_cast<uint64_t*>(smem + a.lds_off_rng) + (size_t)wave * 4 * kRingStride;
    if constexpr (RING) {
#pragma unroll
        for (int G = 0; G < 4; ++G) {
            const int rl = blockIdx.x * 16 + wave * 4 + G;
            uint64_t seed = a.seeds[a.r0 + (rl < a.R ? rl : a.R - 1)];
            if (seed == 0) seed = ~0ull;
            q0[G] = (uint32_t)seed;
            q1[G] = (uint32_t)(seed >> 32);
            q2[G] = 0;
            q3[G] = 0;
        }
        for (int i = 0; i < kRingM * lane; ++i) {
#pragma unroll
            for (int G = 0; G < 4; ++G) {
                const XsStep st = xs_peek(q0[G], q1[G], q2[G], q3[G]);
                q0[G] = q2[G];
                q1[G] = q3[G];
                q2[G] = st.nlo;
                q3[G] = st.nhi;
            }
        }
    } else {
        uint64_t seed = a.seeds[r];
        if (seed == 0) seed = ~0ull;
        alo = (uint32_t)seed;
        ahi = (uint32_t)(seed >> 32);
    }
    // next kRingBatch outputs of read G: lane L writes outputs kRingM * L + 1..
    // of the block, then jumps its sub-stream over the other lanes' outputs
    auto refill = [&](auto gc) {
        constexpr int G = decltype(gc)::value;
        const uint32_t pg = (uint32_t)__builtin_amdgcn_readlane((int)pm, 16 * G);
        uint64_t* rg = ringw + G * kRingStride;
        const uint32_t i0 = (pg & (kRingSize - 1)) + (uint32_t)lane * kRingM;
        uint32_t x0 = q0[G], x1 = q1[G], y0 = q2[G], y1 = q3[G];
        uint64_t o[kRingM];
#pragma unroll
        for (int i = 0; i < kRingM; ++i) {
            const XsStep st = xs_peek(x0, x1, y0, y1);
            o[i] = (uint64_t)st.ohi << 32 | st.olo;
            x0 = y0;
            x1 = y1;
            y0 = st.nlo;
            y1 = st.nhi;
        }
#pragma unroll
        for (int i = 0; i < kRingM; ++i) rg[i0 + i] = o[i];
        if (i0 < 16)  // the first 16 entries are mirrored past the end
#pragma unroll
            for (int i = 0; i < kRingM; ++i) rg[kRingSize + i0 + i] = o[i];
        xs_jump(x0, x1, y0, y1, Jt);
        q0[G] = x0;
        q1[G] = x1;
        q2[G] = y0;
        q3[G] = y1;
        pm += g == G ? (uint32_t)kRingBatch : 0u;
    };
    uint32_t nfl = 0, ndg = 0;
    const int nruns = a.nchain_runs;
    // run bound i: a vector load made uniform (a scalar load would share its
    // wait counter with the LDS operations and return out of order, so every
    // wait for a word would also wait for the atomics issued after it)
    auto bound = [&](int i) {
        int idx = i < nruns ? i : nruns;
        asm("" : "+v"(idx));
        return __builtin_amdgcn_readfirstlane(runs[idx]);
    };

    for (int bi = 0; bi < a.nbetas; ++bi) {
        // this beta's decision table, one copy per wave; every group writes
        // the whole row, so no group depends on another
        uint64_t t0 = 0, t1 = 0, t2 = 0;
        if constexpr (FEW) {
            const uint2* lrow = reinterpret_cast<const uint2*>(a.chain_lut2) + (size_t)bi * a.lut_width;
            for (int t = k; t < a.lut_width; t += 16) Tw2[t] = lrow[t];
            // vector loads made uniform (a pending scalar load would make
            // every LDS wait in the run loop wait for all LDS operations)
            const uint32_t* th = reinterpret_cast<const uint32_t*>(a.chain_thr + (size_t)bi * kChainFewT);
            uint32_t tw[2 * kChainFewT];
#pragma unroll
            for (int i = 0; i < 2 * kChainFewT; ++i) {
                int idx = i;
                asm("" : "+v"(idx));
                tw[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)th[idx]);
            }
            t0 = (uint64_t)tw[1] << 32 | tw[0];
            t1 = (uint64_t)tw[3] << 32 | tw[2];
            t2 = (uint64_t)tw[5] << 32 | tw[4];
        } else {
            const uint4* lrow = reinterpret_cast<const uint4*>(a.chain_lut) + (size_t)bi * a.lut_width;
            for (int t = k; t < a.lut_width; t += 16) Tw[t] = lrow[t];
        }
        __builtin_amdgcn_wave_barrier();
        const uint4* lutw = Tw - a.lut_lo;
        const uint2* lutw2 = Tw2 - a.lut_lo;
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            // run bounds, loaded two runs ahead: the current run is [sA, sB);
            // sC ends the next one (slot records: none)
            int sA = 0, sB = 0, sC = 0;
            if constexpr (!SLOT) {
                sA = bound(0);
                sB = bound(1);
                sC = bound(2);
            }
            // the member's adjacency record, loaded one run ahead; vm: 16 v of
            // its member v (a lane past the run holds the run's first member)
            const uint32_t* srec = SLOT ? a.chain_slots + (size_t)k * kChainRecDwords : nullptr;
            Rec<NR> er;
            uint32_t vm;
            int vmi = 0;  // without slots: the member's index
            if constexpr (SLOT) {
                er = load_rec<NR>(srec, 0);
                vm = er.w[kChainRecDwords - 1] & 0xffff0u;
            } else {
                vmi = k < sB - sA ? sA + k : sA;
                er = load_rec<NR>(rec, vmi);
                vm = (uint32_t)vmi * 16u;
            }
            // the run's word and ring window, loaded during the previous run
            uint32_t wd;
            uint64_t xw[FEW ? 1 : kChainK];
            auto load_next = [&](uint32_t v) {  // v: 16 times the member
                if constexpr (RING) {
                    // a read with fewer than kChainK outputs left gets the next block
                    const uint64_t low = __builtin_amdgcn_ballot_w64(pm - cc < (uint32_t)kChainK);
                    if (low) {
                        if (low & 1ull) refill(std::integral_constant<int, 0>{});
                        if (low & (1ull << 16)) refill(std::integral_constant<int, 1>{});
                        if (low & (1ull << 32)) refill(std::integral_constant<int, 2>{});
                        if (low & (1ull << 48)) refill(std::integral_constant<int, 3>{});
                    }
                }
                wd = *static_cast<const uint32_t*>(__builtin_assume_aligned(smem + (v | pdw), 4));
                if constexpr (FEW) {
                    xw[0] = ringw[g * kRingStride + (cc & (kRingSize - 1)) + wk];
                } else if constexpr (RING) {
                    const uint64_t* win = ringw + g * kRingStride + (cc & (kRingSize - 1));
#pragma unroll
                    for (int i = 0; i < kChainK; ++i) xw[i] = win[i];
                }
            };
            load_next(vm);
            // as many LDS operations after the loads as a run issues after
            // them (its atomics that cannot reach the next run), so that every
            // entry to the loop has the same LDS operations in flight after
            // the word: the wait for it then leaves them out (adds of zero to
            // the member's own word, which no other lane changes meanwhile)
            uint32_t zero = 0u;
            asm("" : "+v"(zero));  // an add the compiler cannot turn into a load
#pragma unroll
            for (int e2 = NCRIT + 1; e2 <= NE; ++e2) lds_add(smem, vm | pdw, zero);
            for (int ri = 0; ri < nruns; ++ri) {
                bool mem;
                uint32_t vn;
                Rec<NR> ern;
                int sD = 0;
                if constexpr (SLOT) {
                    mem = (int32_t)er.w[kChainRecDwords - 1] >= 0;
                    ern = load_rec<NR>(srec, (ri + 1 < nruns ? ri + 1 : ri) * 16);
                    vn = ern.w[kChainRecDwords - 1] & 0xffff0u;
                } else {
                    sD = bound(ri + 3);
                    mem = k < sB - sA;
                    const int vi = sC > sB ? (k < sC - sB ? sB + k : sB) : vmi;
                    ern = load_rec<NR>(rec, vi);
                    vn = (uint32_t)vi * 16u;
                    vmi = vi;
                }

                // 1. both table entries of the member's word
                const int q = (int)((wd >> psh) & 0xffu);
                const int sg = 1 - 2 * (q & 1);
                const int P = (int)(int16_t)(er.w[0] & 0xffffu) * sg;  // successor's patch if this member flips
                int dq = __builtin_amdgcn_update_dpp(0, P, 0x111, 0xf, 0xf, true);  // row_shr:1, lane 0 gets 0
                dq = mem ? dq : 0;
                uint32_t Db;
                if constexpr (FEW) {
                    // entry: shift of its threshold's bits, class bits
                    const uint2 E0 = lutw2[q];
                    const uint2 E1 = lutw2[q + dq];
                    // 2. the group's 15 outputs against each threshold at once:
                    // bit k of the group's 16 is x_{15-k} < T_a
                    const uint64_t b0 = __builtin_amdgcn_ballot_w64(xw[0] < t0);
                    const uint64_t b1 = __builtin_amdgcn_ballot_w64(xw[0] < t1);
                    const uint64_t b2 = __builtin_amdgcn_ballot_w64(xw[0] < t2);
                    const uint32_t s0 = (uint32_t)(b0 >> gsh), s1 = (uint32_t)(b1 >> gsh), s2 = (uint32_t)(b2 >> gsh);
                    const uint64_t X = (uint64_t)(s2 & 0xffffu) << 32 | __builtin_amdgcn_perm(s1, s0, 0x05040100u);
                    const uint32_t m0 = ((uint32_t)(X >> E0.x) & (uint32_t)kChainAlways) | E0.y;
                    const uint32_t m1 = ((uint32_t)(X >> E1.x) & (uint32_t)kChainAlways) | E1.y;
                    Db = m0 | (m1 << 16);
                } else {
                // entry: -T (low, high), T high, class bits (kChainDraw or
                // kChainAlways)
                const uint4 E0 = lutw[q];
                const uint4 E1 = lutw[q + dq];
                const uint64_t nT0 = (uint64_t)E0.y << 32 | E0.x, nT1 = (uint64_t)E1.y << 32 | E1.x;

                // 2. the next kChainK outputs, compared with both entries
                // (bit kChainK - 1 - i: the output after i draws)
                uint32_t m0 = 0, m1 = 0;
                if constexpr (RING) {
#pragma unroll
                    for (int i = 0; i < kChainK; ++i) {
                        m0 = push_less(m0, xw[i], E0.z, nT0);
                        m1 = push_less(m1, xw[i], E1.z, nT1);
                    }
                } else {
                    Xs[0] = (uint64_t)ahi << 32 | alo;
                    Xs[1] = (uint64_t)bhi << 32 | blo;
                    uint32_t plo = alo, phi = ahi, clo = blo, chi = bhi;
#pragma unroll
                    for (int i = 0; i < kChainK; ++i) {
                        const XsStep st = xs_peek(plo, phi, clo, chi);
                        const uint64_t x = (uint64_t)st.ohi << 32 | st.olo;
                        m0 = push_less(m0, x, E0.z, nT0);
                        m1 = push_less(m1, x, E1.z, nT1);
                        Xs[i + 2] = (uint64_t)st.nhi << 32 | st.nlo;
                        plo = clo;
                        phi = chi;
                        clo = st.nlo;
                        chi = st.nhi;
                    }
                }
                Db = (m0 | E0.w) | ((m1 | E1.w) << 16);
                }
                Db = mem ? Db : 0u;

                // 3. serial pass over the members
                uint32_t e = kChainK - 1, pf = 0, flips = 0;
                resolve<0>(Db, e, pf, flips);
                const uint32_t d = (uint32_t)(kChainK - 1) - e;  // draws of the run
                if constexpr (RING) {
                    cc += d;
                } else {
                    const uint64_t na = Xs[d], nb = Xs[d + 1];
                    alo = (uint32_t)na;
                    ahi = (uint32_t)(na >> 32);
                    blo = (uint32_t)nb;
                    bhi = (uint32_t)(nb >> 32);
                }

                // 4. updates of the flipped members: entry (-J) << 24 | u << 4
                // adds (-J) * 4 s to the word of u.  The own bit and the first
                // NCRIT entries (every entry that can name a variable of the
                // next run comes first in the record) go before the next run's
                // loads; the rest after them, overlapping their latency.
                const int fl = (int)((flips >> k) & 1u);  // 0 for lane 15
                const int sf = fl ? sg : 0;
                const int sf4 = sf * 4;
                lds_add(smem, vm | pdw, (uint32_t)sf << psh);
#pragma unroll
                for (int e2 = 1; e2 <= NCRIT; ++e2)
                    lds_add_at(and_or(er.w[e2], 0xffff0u, pdw), (uint32_t)((int)(int8_t)(er.w[e2] >> 24) * sf4) << psh);
                if (ri + 1 < nruns) load_next(vn);
#pragma unroll
                for (int e2 = NCRIT + 1; e2 <= NE; ++e2)
                    lds_add_at(and_or(er.w[e2], 0xffff0u, pdw), (uint32_t)((int)(int8_t)(er.w[e2] >> 24) * sf4) << psh);
                if constexpr (COUNT) {
                    nfl += (uint32_t)fl;
                    ndg += fl ? (er.w[0] >> 16) : 0u;
                }
                vm = vn;
                er = ern;
                sA = sB;
                sB = sC;
                sC = sD;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    if (!act) return;

    // ---- spins out and energy (cpu_sa.cpp:210-227; integers, exact) -------
    for (int v = k; v < n; v += 16) a.states[(size_t)r * n + v] = (int8_t)(1 - 2 * (smem[v * 16 + pbyte] & 1));
    if constexpr (COUNT) {
        for (int o = 8; o >= 1; o >>= 1) {
            nfl += (uint32_t)__shfl_xor((int)nfl, o, 16);
            ndg += (uint32_t)__shfl_xor((int)ndg, o, 16);
        }
    }
    if (k == 0) {
        double e = 0.0;
        for (int v = 0; v < n; ++v) e += (double)(1 - 2 * (smem[v * 16 + pbyte] & 1)) * a.h[v];
        for (int64_t c = 0; c < a.m; ++c) {
            const int su = 1 - 2 * (smem[a.cs[c] * 16 + pbyte] & 1);
            const int sv = 1 - 2 * (smem[a.ce[c] * 16 + pbyte] & 1);
            e += (double)su * a.cw[c] * (double)sv;
        }
        a.energies[r] = e;
        if (a.counters) {
            a.counters[2 * (size_t)r] = nfl;
            a.counters[2 * (size_t)r + 1] = ndg;
        }
    }
}

// The kernel's LDS adds take byte offsets as addresses (lds_add_at): refuse
// to launch a build in which the kernel has static LDS before its dynamic
// array.
template <typename F>
static hipError_t no_static_lds(F fn) {
    hipFuncAttributes at{};
    hipError_t e = hipFuncGetAttributes(&at, reinterpret_cast<const void*>(fn));
    if (e != hipSuccess) return e;
    return at.sharedSizeBytes == 0 ? hipSuccess : hipErrorInvalidDeviceFunction;
}

template <int NE, int NCRIT, bool COUNT, bool RING, bool SLOT>
static hipError_t launch_chain_slot(const SaArgs& a, int nblocks, hipStream_t stream) {
    if constexpr (RING)
        if (a.chain_thr) {
            auto fn = sa_chain_kernel<NE, NCRIT, COUNT, true, true, SLOT>;
            static const hipError_t layout = no_static_lds(fn);
            if (layout != hipSuccess) return layout;
            if (a.lds_bytes > 65536) {
                hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
                if (e != hipSuccess) return e;
            }
            hipLaunchKernelGGL(fn, dim3(nblocks), dim3(256), a.lds_bytes, stream, a, a.chain_rec, a.chain_runs,
                               reinterpret_cast<const uint4*>(a.chain_jump));
            return hipGetLastError();
        }
    auto fn = sa_chain_kernel<NE, NCRIT, COUNT, RING, false, SLOT>;
    static const hipError_t layout = no_static_lds(fn);
    if (layout != hipSuccess) return layout;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(256), a.lds_bytes, stream, a, a.chain_rec, a.chain_runs,
                       reinterpret_cast<const uint4*>(a.chain_jump));
    return hipGetLastError();
}

template <int NE, int NCRIT, bool COUNT, bool RING>
static hipError_t launch_chain_variant(const SaArgs& a, int nblocks, hipStream_t stream) {
    if constexpr (NE <= kChainRecDwords - 2)
        if (a.chain_slots) return launch_chain_slot<NE, NCRIT, COUNT, RING, true>(a, nblocks, stream);
    return launch_chain_slot<NE, NCRIT, COUNT, RING, false>(a, nblocks, stream);
}

template <int NE, bool RING>
static hipError_t launch_chain_ne(const SaArgs& a, int nblocks, hipStream_t stream) {
    // entries that can name a variable of the next run: at most 3 (sorted
    // first; a lattice variable has +1, +L and the +L-1 wrap), or all of them
    if (a.chain_ncrit <= 3)
        return a.counters ? launch_chain_variant<NE, 3, true, RING>(a, nblocks, stream)
                          : launch_chain_variant<NE, 3, false, RING>(a, nblocks, stream);
    return a.counters ? launch_chain_variant<NE, NE, true, RING>(a, nblocks, stream)
                      : launch_chain_variant<NE, NE, false, RING>(a, nblocks, stream);
}

template <bool RING>
static hipError_t launch_chain_ring(const SaArgs& a, int entries, int nblocks, hipStream_t stream) {
    if (entries <= 4) return launch_chain_ne<4, RING>(a, nblocks, stream);
    if (entries <= 6) return launch_chain_ne<6, RING>(a, nblocks, stream);
    if (entries <= 7) return launch_chain_ne<7, RING>(a, nblocks, stream);
    return hipErrorInvalidValue;
}

hipError_t launch_sa_chain(const SaArgs& a, int entries, int nblocks, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    return a.chain_jump ? launch_chain_ring<true>(a, entries, nblocks, stream)
                        : launch_chain_ring<false>(a, entries, nblocks, stream);
}

}  // namespace misa
