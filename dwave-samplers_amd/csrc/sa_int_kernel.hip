// Integer-mode simulated-annealing sweep kernel for gfx950 (the fast path).
//
// Same algorithm and RNG stream as the reference core
// (dwave/samplers/sa/src/cpu_sa.cpp:93-197, xorshift128+ :27-34), for
// problems whose biases are all integers with |h_v| + sum_k |J_vk| <= 63
// (int8 words) or <= 255 (int16 words).
//
// Block = waves_per_block waves; each wave holds 64/L reads, L lanes per read.
// LDS regions (byte offsets in SaArgs):
//   [0, off_rowinfo)  per-read words, variable-major: word (v, p) at
//                     (v << log2 rpb) | p, p the read's permuted slot
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
#include "xorshift.h"

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
__global__ __launch_bounds__(512) void sa_int_kernel(SaArgs a, const uint32_t* __restrict__ rowinfo_g) {
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
    // Word position of this read inside a variable's row of rpb words: the
    // reads of one wave are spread over different dwords, so the atomics of
    // a wave's read groups never target the same dword.
    constexpr int WPD = 4 / (int)sizeof(typename IntWord<ST>::U);  // words per dword
    const int pslot = (a.rpb % WPD == 0) ? (slot % WPD) * (a.rpb / WPD) + slot / WPD : slot;
    const int n = a.n;
    const int rlocal = blockIdx.x * a.rpb + slot;
    const bool act = rlocal < a.R;
    const int r = a.r0 + rlocal;
    const int fmax = a.fmax;

    U* F = reinterpret_cast<U*>(smem);  // word (v, p) at (v << lrpb) | pslot
    uint32_t* RI = reinterpret_cast<uint32_t*>(smem + a.lds_off_rowinfo);
    const uint16_t* ccol = reinterpret_cast<const uint16_t*>(smem + a.lds_off_col);
    const ST* cJ = reinterpret_cast<const ST*>(smem + a.lds_off_J);
    const int tstride = a.twidth > a.lut_width ? a.twidth : a.lut_width;
    uint64_t* Tw = reinterpret_cast<uint64_t*>(smem + a.lds_off_T) + (size_t)wave * tstride;

    // ---- stage row info and the adjacency (coupler insertion order) -------
    for (int v = threadIdx.x; v <= n; v += blockDim.x) RI[v] = a.rowinfo[v];
    if constexpr (CSR_LDS) {
        uint16_t* wcol = reinterpret_cast<uint16_t*>(smem + a.lds_off_col);
        ST* wJ = reinterpret_cast<ST*>(smem + a.lds_off_J);
        for (int64_t k = threadIdx.x; k < a.m2 + 64; k += blockDim.x) {
            wcol[k] = k < a.m2 ? (uint16_t)a.col[k] : (uint16_t)0;
            wJ[k] = k < a.m2 ? (ST)a.Ji[k] : (ST)0;
        }
    }
    // ---- spins in ----------------------------------------------------------
    for (int v = j; v < n; v += L) {
        const int s = act ? (int)a.states[(size_t)r * n + v] : 1;
        F[(v << lrpb) | pslot] = (U)(BIAS + (s < 0 ? 1 : 0));
    }
    __syncthreads();
    // ---- fields h_v + sum_k J_k s_k (integers: order-free) ----------------
    for (int v = j; v < n; v += L) {
        const int kb = a.rowptr[v], ke = a.rowptr[v + 1];
        int e = (int)a.h[v];
        for (int k = kb; k < ke; ++k) {
            const int u = CSR_LDS ? (int)ccol[k] : a.col[k];
            const int Jk = CSR_LDS ? (int)cJ[k] : (int)a.Ji[k];
            e += (1 - 2 * ((int)F[(u << lrpb) | pslot] & 1)) * Jk;
        }
        const int b = (int)F[(v << lrpb) | pslot] & 1;
        F[(v << lrpb) | pslot] = (U)(BIAS + 2 * e + b);
    }
    __syncthreads();
    if (!act) return;

    uint64_t s0 = a.seeds[r];
    if (s0 == 0) s0 = ~0ull;
    uint64_t s1 = 0;
    unsigned long long nflip = 0, ndeg = 0;
    const bool piped = ORDER == 0 && a.pipelined;

    // ---- pipelined sequential loop state (see below) ----------------------
    const uint64_t* lutw = Tw - a.lut_lo;
    const uint32_t lsh = (uint32_t)lrpb + (sizeof(U) == 1 ? 0u : 1u);  // bytes per row of words = 1 << lsh
    const uint32_t pbyte = (uint32_t)pslot * (uint32_t)sizeof(U);
    const uint32_t pdw = pbyte & ~3u;  // this read's dword inside a row (rows are dword aligned)
    const uint32_t psh = (pbyte & 3u) * 8u;
    constexpr uint32_t WMASK = sizeof(U) == 1 ? 0xffu : 0xffffu;
    auto word_dw = [&](int v) -> uint32_t {
        return *static_cast<const uint32_t*>(__builtin_assume_aligned(smem + (((uint32_t)v << lsh) + pdw), 4));
    };
    int q = 0, q1 = 0, ur = 0, Jr = 0;
    uint32_t c0 = 0, c1 = 0, n0 = 0, n1 = 0;  // row info of v (c) and v+1 (n): RI[v], RI[v+1]
    if (piped) {
        q = (int)((word_dw(0) >> psh) & WMASK);
        q1 = (int)((word_dw(1) >> psh) & WMASK);
        c0 = rowinfo_g[0];
        c1 = rowinfo_g[1];
        n0 = rowinfo_g[1];
        n1 = rowinfo_g[2];
        const int rb = (int)(c0 >> 14);
        ur = CSR_LDS ? (int)ccol[rb + j] : a.col[rb + j];  // past the row's end: spare entries
        Jr = CSR_LDS ? (int)cJ[rb + j] : (int)a.Ji[rb + j];
    }

    for (int bi = 0; bi < a.nbetas; ++bi) {
        // this beta's row of the decision table (pipelined loop) or of the
        // acceptance table by dE/2 (simple loop), one copy per wave; every
        // read group writes the whole row, so no group depends on another
        if (piped) {
            const uint64_t* lrow = a.lut + (size_t)bi * a.lut_width;
            for (int t = j; t < a.lut_width; t += L) Tw[t] = lrow[t];
        } else {
            const uint64_t* trow = a.ttab + (size_t)bi * a.twidth;
            for (int t = j; t < a.twidth; t += L) Tw[t] = trow[t];
        }
        __builtin_amdgcn_wave_barrier();
        const int kthr = __builtin_amdgcn_readfirstlane(a.kthr[bi]);
        const uint64_t* T = Tw + fmax;
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            unsigned sweep_flips = 0, sweep_deg = 0;
            if (piped) {
                // Two-ahead pipeline, sequential order, int8 words, n >= 3.
                // At the top of the step for variable v the registers hold:
                //   q, tq  the word of v and its decision-table entry;
                //   q1     the word of v+1 with every flip before v applied;
                //   ur,Jr  this lane's CSR entry of row v; c*/n* row info of v, v+1.
                // The step reads the word of v+2 (before v's own atomics; the
                // flip of v is patched in with j2[v]), the table entries of
                // v+1 for both outcomes of v (patch j1[v]), and the CSR entry
                // of row v+1, all consumed one step later, so no LDS latency
                // lies on the decision chain.  The xorshift step is computed
                // speculatively and committed with selects.  Every lane issues
                // the atomic on every step (no branch, so the count of
                // outstanding LDS operations is static and the end-of-step
                // wait does not cover the atomic): lanes 0..deg-1 add
                // -4 J s_v to their neighbour's word, lane deg adds s_v to
                // v's own word (toggling its spin bit), and the other lanes,
                // or every lane when v does not flip, add 0 to the word of
                // the entry they loaded (a neighbour of a later row, or a
                // spare entry: always a valid word).  When every row fits
                // one group the long-row path is compiled out; when the
                // sweep's counts fit 16 bits the flip and degree counts share
                // one packed register; without a counters buffer they are
                // not kept.
                uint32_t x0lo = (uint32_t)s0, x0hi = (uint32_t)(s0 >> 32);
                uint32_t x1lo = (uint32_t)s1, x1hi = (uint32_t)(s1 >> 32);
                uint64_t tq = lutw[q];
                // wait for this load here, once per sweep: a load still in
                // flight at the loop entry would make every trip of the loop
                // start with a wait for all LDS operations, the previous
                // trip's atomic included
                asm volatile("" : "+v"(tq), "+v"(q), "+v"(q1), "+v"(ur), "+v"(Jr));
                uint32_t acc = 0;            // packed sweep counters: flips (low 16), sum of deg (high 16)
                // FITS: every row fits one group of lanes; PACK: the sweep's
                // flip and degree counts fit 16 bits each; COUNT: keep counts;
                // REG: every row has reg_deg entries (regular graph), so row v
                // starts at v * reg_deg and only the couplings to v+1 and v+2
                // are read from the row info
                const int dreg = a.reg_deg;
                auto step = [&](const int v, const int v2, auto fits_c, auto pack_c, auto count_c, auto reg_c) {
                    constexpr bool FITS = decltype(fits_c)::value;
                    constexpr bool PACK = decltype(pack_c)::value;
                    constexpr bool COUNT = decltype(count_c)::value;
                    constexpr bool REG = decltype(reg_c)::value;
                    // loads consumed by the next step
                    uint32_t w2 = word_dw(v2);
                    uint32_t ri2s = RI[v2];      // row info of v+2 (start, end)
                    uint32_t ri2e = REG ? 0u : RI[v2 + 1];
                    // row v+1 (mod n) starts at rb1
                    const int rb1 = REG ? ((v2 == 0 ? n : v2) - 1) * dreg : (int)(n0 >> 14);
                    // lanes past the row's end read the next rows or the spare entries
                    int ur_n = CSR_LDS ? (int)ccol[rb1 + j] : a.col[rb1 + j];
                    int Jr_n = CSR_LDS ? (int)cJ[rb1 + j] : (int)a.Ji[rb1 + j];
                    const int nbm = __builtin_amdgcn_sbfe(q, 0, 1);   // -b_v
                    const int sg = nbm * 2 + 1;                       // s_v
                    const int m4s = nbm * -8 - 4;                     // -4 s_v
                    const int dq1 = __mul24(m4s, ((int)(c0 << 18)) >> 25);  // -4 J(v,v+1) s_v
                    const int dq2 = __mul24(m4s, ((int)(c0 << 25)) >> 25);  // -4 J(v,v+2) s_v
                    uint64_t t_keep = lutw[q1];
                    uint64_t t_patch = lutw[q1 + dq1];
                    // every load of the step is issued before the rest of its
                    // work, so a whole step lies between a load and its use
                    __builtin_amdgcn_sched_barrier(0);
                    // decision for v: entry 0 rejects and ~0 accepts without a
                    // draw; any other entry draws and accepts iff rand < entry
                    const XsStep st = xs_peek(x0lo, x0hi, x1lo, x1hi);
                    const uint64_t x = (uint64_t)st.ohi << 32 | st.olo;
                    const bool need = (tq + 1ull) > 1ull;
                    const bool flip = (x < tq) | (tq == ~0ull);
                    x0lo = need ? x1lo : x0lo;
                    x0hi = need ? x1hi : x0hi;
                    x1lo = need ? st.nlo : x1lo;
                    x1hi = need ? st.nhi : x1hi;
                    const int rb = REG ? v * dreg : (int)(c0 >> 14);
                    const int deg = REG ? dreg : (int)(c1 >> 14) - rb;
                    {
                        const bool own = j == deg;
                        const uint32_t et = own ? (uint32_t)v : (uint32_t)ur;
                        const int P = own ? sg : __mul24(Jr, m4s);
                        const uint32_t val = (flip && j <= deg) ? ((uint32_t)P << psh) : 0u;
                        uint32_t* w = static_cast<uint32_t*>(__builtin_assume_aligned(smem + ((et << lsh) + pdw), 4));
                        __hip_atomic_fetch_add(w, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                    if (!FITS && deg >= L) {
                        if (flip) {
                            // rows longer than the group: remaining entries, then the own bit
                            for (int k2 = rb + j + L; k2 < rb + deg; k2 += L) {
                                const int u = CSR_LDS ? (int)ccol[k2] : a.col[k2];
                                const int Jk = CSR_LDS ? (int)cJ[k2] : (int)a.Ji[k2];
                                word_add(F, (u << lrpb) | pslot, __mul24(m4s, Jk));
                            }
                            if (j == 0) word_add(F, (v << lrpb) | pslot, sg);
                        }
                    }
                    if constexpr (COUNT) {
                        if constexpr (PACK) {
                            acc += flip ? (((uint32_t)deg << 16) | 1u) : 0u;
                        } else {
                            sweep_flips += flip ? 1u : 0u;
                            sweep_deg += flip ? (unsigned)deg : 0u;
                        }
                    }
                    // advance the pipeline
                    asm volatile("" : "+v"(w2), "+v"(t_keep), "+v"(t_patch), "+v"(ur_n), "+v"(Jr_n), "+v"(ri2s),
                                      "+v"(ri2e));
                    tq = flip ? t_patch : t_keep;
                    q = q1 + (flip ? dq1 : 0);
                    q1 = (int)((w2 >> psh) & WMASK) + (flip ? dq2 : 0);
                    c0 = n0;
                    n0 = __builtin_amdgcn_readfirstlane(ri2s);
                    if constexpr (!REG) {
                        c1 = n1;
                        n1 = __builtin_amdgcn_readfirstlane(ri2e);
                    }
                    ur = ur_n;
                    Jr = Jr_n;
                };
                // two steps per trip: the rotating row-info registers need no copies
                auto sweep = [&](auto fits_c, auto pack_c, auto count_c, auto reg_c) {
                    int it = 0;
                    for (; it + 3 < n; it += 2) {
                        step(it, it + 2, fits_c, pack_c, count_c, reg_c);
                        step(it + 1, it + 3, fits_c, pack_c, count_c, reg_c);
                    }
                    if (it + 2 < n) {
                        step(it, it + 2, fits_c, pack_c, count_c, reg_c);
                        ++it;
                    }
                    step(n - 2, 0, fits_c, pack_c, count_c, reg_c);
                    step(n - 1, 1, fits_c, pack_c, count_c, reg_c);
                };
                using T_ = std::true_type;
                using F_ = std::false_type;
                const bool fits = a.max_row < L;
                const bool pack = n < 65536 && (int64_t)n * a.max_row < 65536;
                const bool reg = fits && dreg >= 0;
                if (a.counters == nullptr) {
                    if (reg) sweep(T_{}, F_{}, F_{}, T_{});
                    else if (fits) sweep(T_{}, F_{}, F_{}, F_{});
                    else sweep(F_{}, F_{}, F_{}, F_{});
                } else if (reg && pack) {
                    sweep(T_{}, T_{}, T_{}, T_{});
                    sweep_flips = acc & 0xffffu;
                    sweep_deg = acc >> 16;
                } else if (fits && pack) {
                    sweep(T_{}, T_{}, T_{}, F_{});
                    sweep_flips = acc & 0xffffu;
                    sweep_deg = acc >> 16;
                } else if (fits) {
                    sweep(T_{}, F_{}, T_{}, F_{});
                } else {
                    sweep(F_{}, F_{}, T_{}, F_{});
                }
                s0 = (uint64_t)x0hi << 32 | x0lo;
                s1 = (uint64_t)x1hi << 32 | x1lo;
            } else {
                for (int it = 0; it < n; ++it) {
                    int v = it;
                    if constexpr (ORDER == 1) {
                        const uint64_t x = xs_next(s0, s1);
                        v = (int)(x % (uint64_t)n);
                    }
                    const int qv = (int)F[(v << lrpb) | pslot];
                    const int rb = (int)(RI[v] >> 14);
                    const int deg = (int)(RI[v + 1] >> 14) - rb;
                    const int b = qv & 1;
                    const int f = (qv >> 1) - BIAS / 2;
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
                            word_add(F, (u << lrpb) | pslot, __mul24(sgn4, Jk));
                        }
                        if (j == 0) word_add(F, (v << lrpb) | pslot, b ? -1 : 1);  // b -> 1 - b
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
        a.states[(size_t)r * n + v] = (int8_t)(1 - 2 * ((int)F[(v << lrpb) | pslot] & 1));
    if (j == 0) {
        double e = 0.0;
        for (int v = 0; v < n; ++v) e += (double)(1 - 2 * ((int)F[(v << lrpb) | pslot] & 1)) * a.h[v];
        for (int64_t c = 0; c < a.m; ++c) {
            const int su = 1 - 2 * ((int)F[(a.cs[c] << lrpb) | pslot] & 1);
            const int sw = 1 - 2 * ((int)F[(a.ce[c] << lrpb) | pslot] & 1);
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
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(64 * a.waves_per_block), a.lds_bytes, stream, a, a.rowinfo);
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
