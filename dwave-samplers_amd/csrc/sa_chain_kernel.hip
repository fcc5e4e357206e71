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
//  2. every lane of the group steps the read's xorshift128+ kChainK times
//     (outputs x_1..x_K, the same in each lane of the group; the state words
//     go to LDS) and compares both entries with every output: bit i of m_h
//     is x_{i+1} < T_h, all ones when T_h = ~0 (flip without a draw), and
//     T_h = 0 (reject without a draw) compares false everywhere;
//  3. a serial pass over the members, each member's bits broadcast along the
//     row (DPP row_newbcast): member k flips iff bit d of m_pf is set, with
//     d the draws so far and pf the flip of member k-1; d grows by one when
//     the member's entry asks for a draw;
//  4. the state after d draws is read back from LDS, and each flipped member
//     adds -4 J s to its neighbours' words and toggles its own spin bit with
//     return-less LDS atomics (the adds commute).
// Words are BIAS + 2 f + b as in sa_int_kernel (b = 1 for spin -1).  Four
// groups per wave, four waves per block (16 reads); word (v, read) sits at
// byte v * 16 + 4 g + w (g the group, w the wave), so the four reads of a
// wave use four different dwords of a row.
#include "sa_engine.h"
#include "xorshift.h"

namespace misa {

namespace {

constexpr int kBias = 128;

__device__ __forceinline__ void lds_add(unsigned char* smem, uint32_t byte, uint32_t val) {
    uint32_t* w = static_cast<uint32_t*>(__builtin_assume_aligned(smem + byte, 4));
    __hip_atomic_fetch_add(w, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Serial pass over the members KK..kChainK-1.  Db (per lane): bits 0..14 m0,
// bit 15 "entry T0 draws", bits 16..30 m1, bit 31 "entry T1 draws"; zero for
// lanes past the run.  d: draws so far, pf: 16 when the previous member
// flipped, flips: bit KK set when member KK flips.
template <int KK>
__device__ __forceinline__ void resolve(uint32_t Db, uint32_t& d, uint32_t& pf, uint32_t& flips) {
    const uint32_t Dk = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)Db, 0x150 + KK, 0xf, 0xf, false);
    const uint32_t fl = (Dk >> (d + pf)) & 1u;
    d += (Dk >> (15u + pf)) & 1u;
    pf = fl << 4;
    flips |= fl << KK;
    if constexpr (KK + 1 < kChainK) resolve<KK + 1>(Db, d, pf, flips);
}

}  // namespace

template <int NE, bool COUNT>
__global__ __launch_bounds__(256) void sa_chain_kernel(SaArgs a, const uint4* __restrict__ rec,
                                                       const int* __restrict__ runs) {
    extern __shared__ __align__(16) unsigned char smem[];
    static_assert(kChainK <= 15, "two 15-bit masks and two draw bits per member fill one dword");
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int k = lane & 15;  // member index within a run
    const int g = lane >> 4;  // group (read) within the wave
    const int slot = wave * 4 + g;
    const int n = a.n;
    const int rlocal = blockIdx.x * 16 + slot;
    const bool act = rlocal < a.R;
    const int r = a.r0 + rlocal;
    const uint32_t pdw = (uint32_t)g * 4u;      // this read's dword in a row of 16 bytes
    const uint32_t psh = (uint32_t)wave * 8u;   // its byte in that dword
    const uint32_t pbyte = pdw + (uint32_t)wave;
    uint64_t* Tw = reinterpret_cast<uint64_t*>(smem + a.lds_off_T) + (size_t)wave * a.lut_width;
    uint64_t* Xs = reinterpret_cast<uint64_t*>(smem + a.lds_off_rng) + (size_t)slot * kChainRngStride;

    // ---- spins in, then fields h_v + sum_k J_k s_k (integers: order-free) --
    for (int v = k; v < n; v += 16) {
        const int s = act ? (int)a.states[(size_t)r * n + v] : 1;
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
    if (!act) return;

    uint64_t seed = a.seeds[r];
    if (seed == 0) seed = ~0ull;
    uint32_t alo = (uint32_t)seed, ahi = (uint32_t)(seed >> 32), blo = 0, bhi = 0;  // (s0, s1)
    uint32_t nfl = 0, ndg = 0;
    const int nruns = a.nchain_runs;
    auto member = [&](int ri) {
        const int v0 = runs[ri], K = runs[ri + 1] - v0;
        return k < K ? v0 + k : v0;
    };

    for (int bi = 0; bi < a.nbetas; ++bi) {
        // this beta's decision table, one copy per wave; every group writes
        // the whole row, so no group depends on another
        const uint64_t* lrow = a.lut + (size_t)bi * a.lut_width;
        for (int t = k; t < a.lut_width; t += 16) Tw[t] = lrow[t];
        __builtin_amdgcn_wave_barrier();
        const uint64_t* lutw = Tw - a.lut_lo;
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            // the member's adjacency record, loaded one run ahead
            int vm = member(0);
            uint4 e0 = rec[2 * vm], e1 = rec[2 * vm + 1];
            for (int ri = 0; ri < nruns; ++ri) {
                const int v0 = runs[ri], K = runs[ri + 1] - v0;
                const bool mem = k < K;
                const int vn = member(ri + 1 < nruns ? ri + 1 : 0);
                const uint4 e0n = rec[2 * vn], e1n = rec[2 * vn + 1];

                // 1. word and both table entries
                const uint32_t wd =
                    *static_cast<const uint32_t*>(__builtin_assume_aligned(smem + ((uint32_t)vm * 16u + pdw), 4));
                const int q = (int)((wd >> psh) & 0xffu);
                const int sg = 1 - 2 * (q & 1);
                const int P = (int)(int16_t)(e1.w & 0xffffu) * sg;  // successor's patch if this member flips
                int dq = __builtin_amdgcn_update_dpp(0, P, 0x111, 0xf, 0xf, true);  // row_shr:1, lane 0 gets 0
                dq = mem ? dq : 0;
                const uint64_t T0 = lutw[q];
                const uint64_t T1 = lutw[q + dq];

                // 2. the next kChainK outputs, compared with both entries
                Xs[0] = (uint64_t)ahi << 32 | alo;
                Xs[1] = (uint64_t)bhi << 32 | blo;
                uint32_t m0 = 0, m1 = 0;
                {
                    uint32_t plo = alo, phi = ahi, clo = blo, chi = bhi;
#pragma unroll
                    for (int i = 0; i < kChainK; ++i) {
                        const XsStep st = xs_peek(plo, phi, clo, chi);
                        const uint64_t x = (uint64_t)st.ohi << 32 | st.olo;
                        m0 |= (uint32_t)(x < T0) << i;
                        m1 |= (uint32_t)(x < T1) << i;
                        Xs[i + 2] = (uint64_t)st.nhi << 32 | st.nlo;
                        plo = clo;
                        phi = chi;
                        clo = st.nlo;
                        chi = st.nhi;
                    }
                }
                constexpr uint32_t kAll = (1u << kChainK) - 1u;
                m0 = T0 == ~0ull ? kAll : m0;
                m1 = T1 == ~0ull ? kAll : m1;
                uint32_t Db = m0 | ((T0 + 1ull) > 1ull ? 1u << 15 : 0u) | (m1 << 16) |
                              ((T1 + 1ull) > 1ull ? 1u << 31 : 0u);
                Db = mem ? Db : 0u;

                // 3. serial pass over the members
                uint32_t d = 0, pf = 0, flips = 0;
                resolve<0>(Db, d, pf, flips);
                const uint64_t na = Xs[d], nb = Xs[d + 1];
                alo = (uint32_t)na;
                ahi = (uint32_t)(na >> 32);
                blo = (uint32_t)nb;
                bhi = (uint32_t)(nb >> 32);

                // 4. updates of the flipped members
                const int fl = (int)((flips >> k) & 1u);
                const int sf = fl ? sg : 0;
                lds_add(smem, (uint32_t)vm * 16u + pdw, (uint32_t)sf << psh);
                const uint32_t ent[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
#pragma unroll
                for (int e = 0; e < NE; ++e) {
                    const uint32_t u = ent[e] >> 16;
                    const int c = (int)(int16_t)(ent[e] & 0xffffu);
                    lds_add(smem, u * 16u + pdw, (uint32_t)(c * sf) << psh);
                }
                if constexpr (COUNT) {
                    nfl += (uint32_t)fl;
                    ndg += fl ? (e1.w >> 16) : 0u;
                }
                vm = vn;
                e0 = e0n;
                e1 = e1n;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();

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

template <int NE, bool COUNT>
static hipError_t launch_chain_variant(const SaArgs& a, const uint32_t* rec, const int* runs, int nblocks,
                                       hipStream_t stream) {
    auto fn = sa_chain_kernel<NE, COUNT>;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(256), a.lds_bytes, stream, a, reinterpret_cast<const uint4*>(rec),
                       runs);
    return hipGetLastError();
}

template <int NE>
static hipError_t launch_chain_ne(const SaArgs& a, const uint32_t* rec, const int* runs, int nblocks,
                                  hipStream_t stream) {
    return a.counters ? launch_chain_variant<NE, true>(a, rec, runs, nblocks, stream)
                      : launch_chain_variant<NE, false>(a, rec, runs, nblocks, stream);
}

hipError_t launch_sa_chain(const SaArgs& a, int entries, int nblocks, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    const uint32_t* rec = a.chain_rec;
    const int* runs = a.chain_runs;
    if (entries <= 4) return launch_chain_ne<4>(a, rec, runs, nblocks, stream);
    if (entries <= 6) return launch_chain_ne<6>(a, rec, runs, nblocks, stream);
    if (entries <= 7) return launch_chain_ne<7>(a, rec, runs, nblocks, stream);
    return hipErrorInvalidValue;
}

}  // namespace misa
