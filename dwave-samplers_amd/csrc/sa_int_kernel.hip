// Integer-mode simulated-annealing sweep kernel for gfx950 (the fast path).
//
// Same algorithm and RNG stream as the reference core
// (dwave/samplers/sa/src/cpu_sa.cpp:93-197, xorshift128+ :27-34), for
// problems whose biases are all integers with |h_v| + sum_k |J_vk| <= 63
// (int8 words) or <= 255 (int16 words).
//
// Block = waves_per_block waves; each wave holds 64/L reads, L lanes per read.
// LDS regions (byte offsets in SaArgs):
//   [0, off_rowinfo)  per-read words, variable-major: word (v, slot) at
//                     (v << log2 rpb) | slot
//   off_rowinfo       (rowptr[v] << 10) | (jnext[v] & 0x3ff), n+1 entries
//   off_col, off_J    CSR column indices (u16) and weights (when staged)
//   off_T             per-wave copy of the current beta's acceptance row
// A word holds BIAS + 2 f_v + b_v (f_v the local field h_v + sum J s, b_v = 1
// for spin -1).  BIAS keeps every word inside its unsigned range, so a flip of
// v adds the state-independent amount -4 J s_v to each neighbour's word with a
// return-less 32-bit LDS atomic on the enclosing dword: no carry or borrow can
// leave the byte (short), and the adds commute, so repeated couplers and
// self-loops need no ordering.  The word of the next variable is loaded one
// step ahead and patched in registers with jnext[v] (the coupling of v to
// v+1), so the decision chain never waits for the atomics.  Needs n >= 2 (the
// next variable is never v itself).  Acceptance uses
// the host-built table T: accept iff rand < T[beta][dE/2] (bit-equal to the
// reference's double comparison for every 64-bit rand).
#include <type_traits>

#include "sa_engine.h"

namespace misa {

namespace {

template <typename ST>
struct IntWord;
template <>
struct IntWord<int8_t> {
    using U = uint8_t;
    static constexpr int BIAS = 128;
};
template <>
struct IntWord<int16_t> {
    using U = uint16_t;
    static constexpr int BIAS = 32768;
};

__device__ __forceinline__ uint64_t xs_next(uint64_t& s0, uint64_t& s1) {
    uint64_t a = s0;
    const uint64_t b = s1;
    s0 = b;
    a ^= a << 23;
    s1 = a ^ b ^ (a >> 17) ^ (b >> 26);
    return s1 + b;
}

// One xorshift128+ step on 32-bit halves, without committing it: returns the
// output and the would-be new state (s0' = s1, s1' = n).
struct XsStep {
    uint32_t nlo, nhi;   // new s1
    uint32_t olo, ohi;   // output n + s1
};
__device__ __forceinline__ XsStep xs_peek(uint32_t alo, uint32_t ahi, uint32_t blo, uint32_t bhi) {
    // a ^= a << 23
    const uint32_t tlo = alo << 23;
    const uint32_t thi = __builtin_amdgcn_alignbit(ahi, alo, 9);
    alo ^= tlo;
    ahi ^= thi;
    // n = a ^ b ^ (a >> 17) ^ (b >> 26)
    XsStep r;
    r.nlo = alo ^ blo ^ __builtin_amdgcn_alignbit(ahi, alo, 17) ^ __builtin_amdgcn_alignbit(bhi, blo, 26);
    r.nhi = ahi ^ bhi ^ (ahi >> 17) ^ (bhi >> 26);
    const uint64_t o = ((uint64_t)r.nhi << 32 | r.nlo) + ((uint64_t)bhi << 32 | blo);
    r.olo = (uint32_t)o;
    r.ohi = (uint32_t)(o >> 32);
    return r;
}

// Adds `delta` to the element at index e of an array of U in LDS through a
// return-less atomic add on the aligned dword holding it.
template <typename U>
__device__ __forceinline__ void word_add(U* base, int e, int delta) {
    const uint32_t byteoff = (uint32_t)e * (uint32_t)sizeof(U);
    uint32_t* w = reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(base) + (byteoff & ~3u));
    const uint32_t sh = (byteoff & 3u) * 8u;
    __hip_atomic_fetch_add(w, (uint32_t)delta << sh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

}  // namespace

template <typename ST, int ORDER, int PROP, bool CSR_LDS>
__global__ __launch_bounds__(256) void sa_int_kernel(SaArgs a) {
    using U = typename IntWord<ST>::U;
    constexpr int BIAS = IntWord<ST>::BIAS;
    extern __shared__ __align__(16) unsigned char smem[];

    const int lg = a.lanes_log2;
    const int L = 1 << lg;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & (L - 1);
    const int rpw = 64 >> lg;
    const int slot = wave * rpw + (lane >> lg);
    const int lrpb = a.rpb_log2;
    const int n = a.n;
    const int rlocal = blockIdx.x * a.rpb + slot;
    const bool act = rlocal < a.R;
    const int r = a.r0 + rlocal;
    const int fmax = a.fmax;

    U* F = reinterpret_cast<U*>(smem);  // word (v, slot) at (v << lrpb) | slot
    uint32_t* RI = reinterpret_cast<uint32_t*>(smem + a.lds_off_rowinfo);
    const uint16_t* ccol = reinterpret_cast<const uint16_t*>(smem + a.lds_off_col);
    const ST* cJ = reinterpret_cast<const ST*>(smem + a.lds_off_J);
    uint64_t* Tw = reinterpret_cast<uint64_t*>(smem + a.lds_off_T) + (size_t)wave * a.twidth;

    // ---- stage row info and the adjacency (coupler insertion order) -------
    for (int v = threadIdx.x; v <= n; v += blockDim.x) RI[v] = a.rowinfo[v];
    if constexpr (CSR_LDS) {
        uint16_t* wcol = reinterpret_cast<uint16_t*>(smem + a.lds_off_col);
        ST* wJ = reinterpret_cast<ST*>(smem + a.lds_off_J);
        for (int64_t k = threadIdx.x; k < a.m2; k += blockDim.x) {
            wcol[k] = (uint16_t)a.col[k];
            wJ[k] = (ST)a.Ji[k];
        }
    }
    // ---- spins in ----------------------------------------------------------
    for (int v = j; v < n; v += L) {
        const int s = act ? (int)a.states[(size_t)r * n + v] : 1;
        F[(v << lrpb) | slot] = (U)(BIAS + (s < 0 ? 1 : 0));
    }
    __syncthreads();
    // ---- fields h_v + sum_k J_k s_k (integers: order-free) ----------------
    for (int v = j; v < n; v += L) {
        const int kb = a.rowptr[v], ke = a.rowptr[v + 1];
        int e = (int)a.h[v];
        for (int k = kb; k < ke; ++k) {
            const int u = CSR_LDS ? (int)ccol[k] : a.col[k];
            const int Jk = CSR_LDS ? (int)cJ[k] : (int)a.Ji[k];
            e += (1 - 2 * ((int)F[(u << lrpb) | slot] & 1)) * Jk;
        }
        const int b = (int)F[(v << lrpb) | slot] & 1;
        F[(v << lrpb) | slot] = (U)(BIAS + 2 * e + b);
    }
    __syncthreads();
    if (!act) return;

    uint64_t s0 = a.seeds[r];
    if (s0 == 0) s0 = ~0ull;
    uint64_t s1 = 0;
    unsigned long long nflip = 0, ndeg = 0;

    // Word of variable 0 and its row info (sequential order runs one ahead).
    int q = (int)F[slot];
    uint32_t ri0 = RI[0], ri1 = RI[n > 0 ? 1 : 0];

    for (int bi = 0; bi < a.nbetas; ++bi) {
        // acceptance row of this beta, one copy per wave (every read group
        // writes the whole row, so no group depends on another)
        const uint64_t* trow = a.ttab + (size_t)bi * a.twidth;
        for (int t = j; t < a.twidth; t += L) Tw[t] = trow[t];
        __builtin_amdgcn_wave_barrier();
        const int kthr = __builtin_amdgcn_readfirstlane(a.kthr[bi]);
        const uint64_t* T = Tw + fmax;
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            unsigned sweep_flips = 0, sweep_deg = 0;
            if constexpr (ORDER == 0) {
                // Sequential order: branch-free decision, one exec region for
                // the neighbour atomics.  Row bounds and the words of variable
                // it+1 are loaded one step ahead.
                uint32_t x0lo = (uint32_t)s0, x0hi = (uint32_t)(s0 >> 32);
                uint32_t x1lo = (uint32_t)s1, x1hi = (uint32_t)(s1 >> 32);
                uint32_t r0w = __builtin_amdgcn_readfirstlane(ri0);
                uint32_t r1w = __builtin_amdgcn_readfirstlane(ri1);
                for (int it = 0; it < n; ++it) {
                    const int v = it;
                    const int vn = (it + 1 == n) ? 0 : it + 1;
                    const uint32_t byteoff = (uint32_t)((vn << lrpb) | slot) * (uint32_t)sizeof(U);
                    uint32_t qdw = *reinterpret_cast<const uint32_t*>(smem + (byteoff & ~3u));
                    const uint32_t qsh = (byteoff & 3u) * 8u;
                    uint32_t ri0n = RI[vn];
                    uint32_t ri1n = RI[vn + 1];
                    const int rb = (int)(r0w >> 10);
                    const int deg = (int)(r1w >> 10) - rb;
                    // this lane's CSR entry (clamped into the row; used only on a flip)
                    const int kk = rb + (j < deg ? j : 0);
                    int ur = CSR_LDS ? (int)ccol[kk] : a.col[kk];
                    int Jr = CSR_LDS ? (int)cJ[kk] : (int)a.Ji[kk];
                    const int b = q & 1;
                    const int f = (q >> 1) - BIAS / 2;
                    const int k = b ? f : -f;  // dE / 2, |k| <= fmax
                    const uint64_t t = T[k];
                    const XsStep st = xs_peek(x0lo, x0hi, x1lo, x1hi);
                    const bool live = k < kthr;
                    const bool need = PROP == 1 ? (live && k > 0) : live;
                    const bool down = PROP == 1 ? (live && k <= 0) : false;
                    const uint64_t x = (uint64_t)st.ohi << 32 | st.olo;
                    const bool flip = down || (need && x < t);
                    if (need) {
                        x0lo = x1lo;
                        x0hi = x1hi;
                        x1lo = st.nlo;
                        x1hi = st.nhi;
                    }
                    asm volatile("" : "+v"(ur), "+v"(Jr));
                    if (flip && j <= deg) {
                        // neighbour j (j < deg) gets -4 J s_v; lane deg toggles
                        // the spin bit of v (+s_v)
                        const bool nb = j < deg;
                        const int et = ((nb ? ur : v) << lrpb) | slot;
                        const int P = nb ? 4 * Jr : -1;
                        word_add(F, et, b ? P : -P);
                    }
                    if (deg >= L && flip) {
                        // rows longer than the group: remaining entries, then the own bit
                        const int sgn4 = b ? 4 : -4;
                        for (int k2 = rb + j + L; k2 < rb + deg; k2 += L) {
                            const int u = CSR_LDS ? (int)ccol[k2] : a.col[k2];
                            const int Jk = CSR_LDS ? (int)cJ[k2] : (int)a.Ji[k2];
                            word_add(F, (u << lrpb) | slot, sgn4 * Jk);
                        }
                        if (j == 0) word_add(F, (v << lrpb) | slot, b ? -1 : 1);
                    }
                    const int jn4 = 4 * (((int)(r0w << 22)) >> 22);  // 4 * jnext[v]
                    const int qpatch = flip ? (b ? jn4 : -jn4) : 0;
                    sweep_flips += flip ? 1u : 0u;
                    sweep_deg += flip ? (unsigned)deg : 0u;
                    // consume the prefetched values only here, after the decision
                    asm volatile("" : "+v"(qdw), "+v"(ri0n), "+v"(ri1n));
                    q = (int)((qdw >> qsh) & (sizeof(U) == 1 ? 0xffu : 0xffffu)) + qpatch;
                    r0w = __builtin_amdgcn_readfirstlane(ri0n);
                    r1w = __builtin_amdgcn_readfirstlane(ri1n);
                }
                ri0 = r0w;
                ri1 = r1w;
                s0 = (uint64_t)x0hi << 32 | x0lo;
                s1 = (uint64_t)x1hi << 32 | x1lo;
            } else {
            for (int it = 0; it < n; ++it) {
                int v;
                {
                    const uint64_t x = xs_next(s0, s1);
                    v = (int)(x % (uint64_t)n);
                    q = (int)F[(v << lrpb) | slot];
                    ri0 = RI[v];
                    ri1 = RI[v + 1];
                }
                const int rb = (int)(ri0 >> 10);
                const int deg = (int)(ri1 >> 10) - rb;
                const int b = q & 1;
                const int f = (q >> 1) - BIAS / 2;
                const int k = b ? f : -f;  // dE / 2
                bool flip = false;
                if (k < kthr) {
                    if constexpr (PROP == 1) {
                        if (k <= 0) {
                            flip = true;
                        } else {
                            const uint64_t t = T[k];
                            const uint64_t x = xs_next(s0, s1);
                            flip = x < t;
                        }
                    } else {
                        const uint64_t t = T[k];
                        const uint64_t x = xs_next(s0, s1);
                        flip = x < t;
                    }
                }
                if (flip) {
                    const int sgn4 = b ? 4 : -4;
                    for (int k2 = rb + j; k2 < rb + deg; k2 += L) {
                        const int u = CSR_LDS ? (int)ccol[k2] : a.col[k2];
                        const int Jk = CSR_LDS ? (int)cJ[k2] : (int)a.Ji[k2];
                        word_add(F, (u << lrpb) | slot, sgn4 * Jk);
                    }
                    if (j == 0) word_add(F, (v << lrpb) | slot, b ? -1 : 1);  // b -> 1 - b
                    sweep_flips += 1;
                    sweep_deg += (unsigned)deg;
                }
            }
            }
            nflip += sweep_flips;
            ndeg += sweep_deg;
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- spins out and energy (cpu_sa.cpp:210-227; integers, exact) -------
    for (int v = j; v < n; v += L)
        a.states[(size_t)r * n + v] = (int8_t)(1 - 2 * ((int)F[(v << lrpb) | slot] & 1));
    if (j == 0) {
        double e = 0.0;
        for (int v = 0; v < n; ++v) e += (double)(1 - 2 * ((int)F[(v << lrpb) | slot] & 1)) * a.h[v];
        for (int64_t c = 0; c < a.m; ++c) {
            const int su = 1 - 2 * ((int)F[(a.cs[c] << lrpb) | slot] & 1);
            const int sw = 1 - 2 * ((int)F[(a.ce[c] << lrpb) | slot] & 1);
            e += (double)su * a.cw[c] * (double)sw;
        }
        a.energies[r] = e;
        if (a.counters) {
            a.counters[2 * (size_t)r] = nflip;
            a.counters[2 * (size_t)r + 1] = ndeg;
        }
    }
}

template <typename ST, int ORDER, int PROP, bool CSR>
static hipError_t launch_int_variant(const SaArgs& a, int nblocks, hipStream_t stream) {
    auto fn = sa_int_kernel<ST, ORDER, PROP, CSR>;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(64 * a.waves_per_block), a.lds_bytes, stream, a);
    return hipGetLastError();
}

template <typename ST, int ORDER, int PROP>
static hipError_t launch_int_csr(const SaArgs& a, bool csr, int nb, hipStream_t s) {
    return csr ? launch_int_variant<ST, ORDER, PROP, true>(a, nb, s)
               : launch_int_variant<ST, ORDER, PROP, false>(a, nb, s);
}

template <typename ST>
static hipError_t launch_int_mode(const SaArgs& a, int order, int prop, bool csr, int nb, hipStream_t s) {
    if (order == 0)
        return prop == 1 ? launch_int_csr<ST, 0, 1>(a, csr, nb, s) : launch_int_csr<ST, 0, 0>(a, csr, nb, s);
    return prop == 1 ? launch_int_csr<ST, 1, 1>(a, csr, nb, s) : launch_int_csr<ST, 1, 0>(a, csr, nb, s);
}

hipError_t launch_sa_int_lds(const SaArgs& a, int mode, int order, int proposal, bool csr_in_lds, int nblocks,
                             hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    if (mode == MODE_INT8) return launch_int_mode<int8_t>(a, order, proposal, csr_in_lds, nblocks, stream);
    return launch_int_mode<int16_t>(a, order, proposal, csr_in_lds, nblocks, stream);
}

}  // namespace misa
