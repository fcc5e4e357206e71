// The per-block body of the float64 simulated-annealing sweep (shared by the
// plain launch in sa_real_kernel.hip and the batched launch in
// sa_seg_kernel.hip).  Algorithm, layout and pipelining are described in
// sa_real_kernel.hip.
#pragma once

#include <type_traits>

#include "sa_engine.h"

namespace misa {

namespace real_detail {

__device__ __forceinline__ uint64_t xs_step_r(uint64_t& s0, uint64_t& s1) {
    uint64_t a = s0;
    const uint64_t b = s1;
    s0 = b;
    a ^= a << 23;
    s1 = a ^ b ^ (a >> 17) ^ (b >> 26);
    return s1 + b;
}

constexpr double kTwo64 = 18446744073709551616.0;  // (double)RANDMAX, cpu_sa.cpp:154,163

// The acceptance tests of cpu_sa.cpp:154 and :163, settled in single
// precision when a proven margin allows, else by the reference's double
// expression.  e = exp2f(t log2 e) (v_exp_f32, about 1 ulp): for |t| <= 87
// the float conversions of t and of t log2 e cost at most 1.5e-5 in the
// exponent, so e, and 1 + e, are within a relative 1.2e-5 of exp(t) and
// 1 + exp(t); the float products add 2^-24 each.  With xh = x >> 40, x lies in
// [xh 2^40, (xh + 1) 2^40), and a relative margin of 2^-14 (6.1e-5) covers all
// of it with room for the double expression's own roundings (3 2^-53).
//
// Metropolis, t = -dE beta < 0: exp(t) 2^64 > (double)x.  Below t = -87, e is
// zero or subnormal while exp(t) 2^24 < 2^-101, so e 2^24 (1 + 2^-14) < xh
// still proves rejection whenever xh >= 1.
__device__ __forceinline__ bool metropolis_accept(const double t, const uint64_t x) {
    const float xh = (float)(uint32_t)(x >> 40);
    const float a = __builtin_amdgcn_exp2f((float)t * 1.44269504f) * 16777216.0f;
    if (a * 0.99993896f > xh + 1.0f) return true;
    if (a * 1.00006104f < xh) return false;
    return exp(t) * kTwo64 > (double)x;
}

// Gibbs, t = dE beta < 44.37 (draws happen below the threshold 44.36142 /
// beta): 2^64 > (double)x (1 + exp(t)).  (xh + 1) (1 + e) (1 + 2^-14) < 2^24
// proves the flip, xh (1 + e) (1 - 2^-14) > 2^24 proves none (both products
// stay below 2^24 * 2^65, far from float overflow).  Below t = -87, 1 + e and
// 1 + exp(t) both round to 1 within 2^-101.
__device__ __forceinline__ bool gibbs_accept(const double t, const uint64_t x) {
    const float xh = (float)(uint32_t)(x >> 40);
    const float g = 1.0f + __builtin_amdgcn_exp2f((float)t * 1.44269504f);
    if ((xh + 1.0f) * g * 1.00006104f < 16777216.0f) return true;
    if (xh * g * 0.99993896f > 16777216.0f) return false;
    return kTwo64 > (double)x * (1.0 + exp(t));
}

}  // namespace real_detail

// One block's share of a launch: reads blk*rpb ... blk*rpb + rpb - 1 of args.
template <int ORDER, int PROP, bool CSR_LDS>
__device__ __forceinline__ void real_block(const SaArgs& a, const int blk, unsigned char* smem) {
    const int lg = a.lanes_log2;
    const int L = 1 << lg;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & (L - 1);
    const int rpw = 64 >> lg;
    const int slot = wave * rpw + (lane >> lg);
    const int lrpb = a.rpb_log2;
    const int n = a.n;
    const int rlocal = blk * a.rpb + slot;
    // slots past rpb exist when a block has more waves than reads (batched
    // layouts of large problems); they only help stage the adjacency
    const bool own = slot < a.rpb;
    const bool act = own && rlocal < a.R;
    const int r = a.r0 + rlocal;

    double* F = reinterpret_cast<double*>(smem);
    int8_t* S = reinterpret_cast<int8_t*>(smem + a.lds_off_J_spins);
    uint32_t* RI = reinterpret_cast<uint32_t*>(smem + a.lds_off_rowinfo);
    const uint16_t* ccol = reinterpret_cast<const uint16_t*>(smem + a.lds_off_col);
    const double* cJ = reinterpret_cast<const double*>(smem + a.lds_off_J);

    // ---- stage row info and the adjacency (coupler insertion order) -------
    for (int v = threadIdx.x; v <= n; v += blockDim.x)
        RI[v] = (uint32_t)a.rowptr[v] | (v < n && a.serial_row[v] ? 0x80000000u : 0u) |
                (v < n && a.next_adj[v] ? 0x40000000u : 0u);
    if constexpr (CSR_LDS) {
        uint16_t* wcol = reinterpret_cast<uint16_t*>(smem + a.lds_off_col);
        double* wJ = reinterpret_cast<double*>(smem + a.lds_off_J);
        // 64 spare entries past the last row (lanes past a row's end read them)
        for (int64_t k = threadIdx.x; k < a.m2 + 64; k += blockDim.x) {
            wcol[k] = k < a.m2 ? (uint16_t)a.col[k] : (uint16_t)0;
            wJ[k] = k < a.m2 ? a.Jd[k] : 0.0;
        }
    }
    // ---- spins in ----------------------------------------------------------
    if (own)
        for (int v = j; v < n; v += L)
            S[(v << lrpb) | slot] = act ? a.states[(size_t)r * n + v] : (int8_t)1;
    __syncthreads();
    // ---- fields, summed in adjacency order (cpu_sa.cpp:62-68) --------------
    if (own)
    for (int v = j; v < n; v += L) {
        const int kb = a.rowptr[v], ke = a.rowptr[v + 1];
        double e = a.h[v];
        for (int k = kb; k < ke; ++k) {
            const int u = CSR_LDS ? (int)ccol[k] : a.col[k];
            const double Jk = CSR_LDS ? cJ[k] : a.Jd[k];
            e += (double)S[(u << lrpb) | slot] * Jk;
        }
        F[(v << lrpb) | slot] = e;
    }
    __syncthreads();
    if (!act) return;

    uint64_t s0 = a.seeds[r];
    if (s0 == 0) s0 = ~0ull;
    uint64_t s1 = 0;
    unsigned long long nflip = 0, ndeg = 0;

    // cpu_sa.cpp:140-166: skip above the threshold, else Metropolis / Gibbs
    auto decide = [&](const double f, const int sv, const double beta, const double thr) -> bool {
        const double de = (double)(-2 * sv) * f;
        bool flip = false;
        if (!(de >= thr)) {  // cpu_sa.cpp:140
            if constexpr (PROP == 1) {
                if (de <= 0.0) {
                    flip = true;
                } else {
                    const uint64_t x = real_detail::xs_step_r(s0, s1);
                    flip = real_detail::metropolis_accept(-de * beta, x);  // :154
                }
            } else {
                const uint64_t x = real_detail::xs_step_r(s0, s1);
                flip = real_detail::gibbs_accept(de * beta, x);  // :163
            }
        }
        return flip;
    };

    bool pipelined = false;
    // CSR entry k, staged in LDS or in device memory
    auto col_at = [&](const int k) -> int { return CSR_LDS ? (int)ccol[k] : a.col[k]; };
    auto w_at = [&](const int k) -> double { return CSR_LDS ? cJ[k] : a.Jd[k]; };
    if constexpr (ORDER == 0) {
        // with the adjacency in LDS the pipeline pays off when every row
        // fits one pass of the lanes (measured: long rows run faster unpipelined)
        if (n >= 2 && (!CSR_LDS || a.max_row < L)) {
            pipelined = true;
            constexpr uint32_t ROW = 0x3fffffffu, SERIAL = 0x80000000u, ADJ = 0x40000000u;
            // rows v (A), v+1 (B), v+2 (C), mod n: row info (start word, end word)
            uint32_t aS = RI[0], aE = RI[1], bS = RI[1], bE = RI[2];
            const int w0 = n > 2 ? 2 : 0;
            uint32_t cS = RI[w0], cE = RI[w0 + 1];
            // this lane's CSR entry and J(w, w+1) of rows A and B (the arrays
            // carry 64 spare entries past the last row: every load is in bounds)
            int urA = col_at((int)(aS & ROW) + j), urB = col_at((int)(bS & ROW) + j);
            double JrA = w_at((int)(aS & ROW) + j), JrB = w_at((int)(bS & ROW) + j);
            double jnA = a.jnext_d[0], jnB = a.jnext_d[1];
            double fC = F[slot];  // field and spin of v
            int sC = S[slot];
            for (int bi = 0; bi < a.nbetas; ++bi) {
                double beta = a.betas[bi];
                double thr = a.thr[bi];  // 44.36142 / beta, cpu_sa.cpp:131
                // landed here: the steps' waits for their own loads then count
                // no older device-memory load
                asm volatile("" : "+v"(beta), "+v"(thr));
                for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
                    unsigned sweep_flips = 0, sweep_deg = 0;
                    auto step = [&](const int v) {
                        const int nx = v + 1 < n ? v + 1 : 0;
                        const int w2 = nx + 1 < n ? nx + 1 : 0;
                        const int w3 = w2 + 1 < n ? w2 + 1 : 0;
                        // loads for later steps, issued before this step's updates
                        double fN = F[(nx << lrpb) | slot];
                        int sN = S[(nx << lrpb) | slot];
                        uint32_t dS = RI[w3], dE = RI[w3 + 1];
                        const int rbC = (int)(cS & ROW);
                        int urC = col_at(rbC + j);
                        double JrC = w_at(rbC + j);
                        const double jnC = a.jnext_d[w2];
                        __builtin_amdgcn_sched_barrier(0);
                        const bool flip = decide(fC, sC, beta, thr);
                        asm volatile("" : "+v"(fN), "+v"(sN), "+v"(dS), "+v"(dE));
                        if constexpr (CSR_LDS) asm volatile("" : "+v"(urC), "+v"(JrC));
                        if (flip) {
                            const int e = (v << lrpb) | slot;
                            const int rb = (int)(aS & ROW);
                            const int deg = (int)(aE & ROW) - rb;
                            const double m2s = (double)(-2 * sC);  // field change per unit J: -2 s_v
                            if (aS & SERIAL) {
                                // repeated neighbours / self-loops: one lane, CSR order
                                if (j == 0) {
                                    for (int k = rb; k < rb + deg; ++k)
                                        F[(col_at(k) << lrpb) | slot] += m2s * w_at(k);
                                    S[e] = (int8_t)(-sC);
                                }
                                fN = F[(nx << lrpb) | slot];  // after the ordered updates
                                asm volatile("" : "+v"(fN));  // landed here, not at the next step's top
                            } else {
                                if (j < deg)
                                    __hip_atomic_fetch_add(&F[(urA << lrpb) | slot], m2s * JrA, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                                for (int k = rb + j + L; k < rb + deg; k += L)
                                    __hip_atomic_fetch_add(&F[(col_at(k) << lrpb) | slot], m2s * w_at(k),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                if (j == 0) S[e] = (int8_t)(-sC);
                                if (aS & ADJ) fN = fN + m2s * jnA;
                            }
                            sweep_flips += 1;
                            sweep_deg += (unsigned)deg;
                        }
                        fC = fN;
                        sC = sN;
                        aS = bS; aE = bE; bS = cS; bE = cE; cS = dS; cE = dE;
                        urA = urB; JrA = JrB; jnA = jnB;
                        urB = urC; JrB = JrC; jnB = jnC;
                        __builtin_amdgcn_wave_barrier();
                    };
                    // two steps per trip: the device-memory loads of a trip
                    // become rows A and B of the next without register copies
                    // (a copy would wait for the load)
                    int v = 0;
                    for (; v + 1 < n; v += 2) {
                        step(v);
                        step(v + 1);
                    }
                    if (v < n) step(v);
                    nflip += sweep_flips;
                    ndeg += sweep_deg;
                }
            }
        }
    }

    for (int bi = 0; bi < (pipelined ? 0 : a.nbetas); ++bi) {
        const double beta = a.betas[bi];
        const double thr = a.thr[bi];  // 44.36142 / beta, cpu_sa.cpp:131
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            unsigned sweep_flips = 0, sweep_deg = 0;
            for (int it = 0; it < n; ++it) {
                int v = it;
                if constexpr (ORDER == 1) {
                    const uint64_t x = real_detail::xs_step_r(s0, s1);
                    v = (int)(x % (uint64_t)n);
                }
                const int e = (v << lrpb) | slot;
                const double f = F[e];
                const int sv = S[e];
                const uint32_t r0 = RI[v], r1 = RI[v + 1];
                const int rb = (int)(r0 & 0x3fffffffu);
                const int deg = (int)(r1 & 0x3fffffffu) - rb;
                // this lane's CSR entry (used only on a flip)
                int ur = 0;
                double Jr = 0.0;
                if (j < deg) {
                    ur = CSR_LDS ? (int)ccol[rb + j] : a.col[rb + j];
                    Jr = CSR_LDS ? cJ[rb + j] : a.Jd[rb + j];
                }
                const double de = (double)(-2 * sv) * f;
                bool flip = false;
                if (!(de >= thr)) {  // cpu_sa.cpp:140
                    if constexpr (PROP == 1) {
                        if (de <= 0.0) {
                            flip = true;
                        } else {
                            const uint64_t x = real_detail::xs_step_r(s0, s1);
                            flip = real_detail::metropolis_accept(-de * beta, x);  // :154
                        }
                    } else {
                        const uint64_t x = real_detail::xs_step_r(s0, s1);
                        flip = real_detail::gibbs_accept(de * beta, x);  // :163
                    }
                }
                if (flip) {
                    const double m2s = (double)(-2 * sv);  // field change per unit J: -2 s_v
                    if (r0 & 0x80000000u) {
                        // repeated neighbours / self-loops: one lane, CSR order
                        if (j == 0) {
                            for (int k = rb; k < rb + deg; ++k) {
                                const int u = CSR_LDS ? (int)ccol[k] : a.col[k];
                                const double Jk = CSR_LDS ? cJ[k] : a.Jd[k];
                                F[(u << lrpb) | slot] += m2s * Jk;
                            }
                            S[e] = (int8_t)(-sv);
                        }
                    } else {
                        if (j < deg)
                            __hip_atomic_fetch_add(&F[(ur << lrpb) | slot], m2s * Jr, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        for (int k = rb + j + L; k < rb + deg; k += L) {
                            const int u = CSR_LDS ? (int)ccol[k] : a.col[k];
                            const double Jk = CSR_LDS ? cJ[k] : a.Jd[k];
                            __hip_atomic_fetch_add(&F[(u << lrpb) | slot], m2s * Jk, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                        if (j == 0) S[e] = (int8_t)(-sv);
                    }
                    sweep_flips += 1;
                    sweep_deg += (unsigned)deg;
                }
                __builtin_amdgcn_wave_barrier();
            }
            nflip += sweep_flips;
            ndeg += sweep_deg;
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- spins out and energy (cpu_sa.cpp:210-227) -------------------------
    for (int v = j; v < n; v += L) a.states[(size_t)r * n + v] = S[(v << lrpb) | slot];
    if (j == 0) {
        double en = 0.0;
        for (int v = 0; v < n; ++v) en += (double)S[(v << lrpb) | slot] * a.h[v];
        for (int64_t c = 0; c < a.m; ++c)
            en += (double)S[(a.cs[c] << lrpb) | slot] * a.cw[c] * (double)S[(a.ce[c] << lrpb) | slot];
        a.energies[r] = en;
        if (a.counters) {
            a.counters[2 * (size_t)r] = nflip;
            a.counters[2 * (size_t)r + 1] = ndeg;
        }
    }
}

}  // namespace misa
