// Dense float64 sweep kernel for small problems (n <= 128, sequential order,
// no repeated couplers or self-loops) on gfx950.
//
// Same algorithm, float64 operation order and RNG stream as the reference
// core (dwave/samplers/sa/src/cpu_sa.cpp:93-197, xorshift128+ :27-34) and as
// sa_real_kernel; what differs is where the state lives.  One wave holds one
// read:
//  * the fields f_v (float64) in VGPRs: variable 64 i + l in register i of
//    lane l (NR registers);
//  * the spins as NR 64-bit masks and the xorshift128+ state in SGPRs;
//  * the couplings as a dense matrix in LDS, row v holding J(v, u) at u (zero
//    where v and u are not coupled), shared by the block's waves.
// A step reads f_v with v_readlane, so the decision is uniform over the wave
// and its branches are scalar; an accepted flip of v updates every field at
// once, f_u = fma(-2 s_v, J(v, u), f_u) in each lane.  -2 s_v J is exact, so
// the fused form rounds exactly like the reference's dE_u += 4 s_v J s_u; for
// the zero entries it adds a signed zero, which can change only the sign of a
// zero field, and no decision depends on that sign (dE = +-0 compares equal,
// exp(+-0) = 1, and energies are recomputed from the spins).
//
// Acceptance of a draw is settled first in single precision: for Metropolis,
// flip iff exp(t) 2^64 > (double)x (t = -dE beta in (-44.37, 0), x the draw,
// cpu_sa.cpp:154).  a = exp2f(t log2 e) 2^24 is within 8.2e-6 (relative) of
// exp(t) 2^24: (float)t and the product with log2 e cost at most 1.14e-5 in
// the exponent (|t log2 e| < 64), v_exp_f32 about 1 ulp.  With xh = x >> 40,
// x lies in [xh 2^40, (xh + 1) 2^40), so a (1 - 2^-14) > xh + 1 proves the
// flip and a (1 + 2^-14) < xh proves the rejection, whatever the roundings of
// the double expression (about 2^-52); only draws within about 6e-5 of the
// acceptance probability evaluate the reference's double expression.  Gibbs
// (flip iff 2^64 > (double)x (1 + exp(dE beta)), :163) uses
// a = 2^24 rcp(1 + exp2f(dE beta log2 e)) with the same margins (its error is
// that of 1 + exp2f, at most the same relative bound, plus 1 ulp of v_rcp_f32).
//
// The next variable's field is read before the step's update and patched
// with the same fma on the same operands when v flips, so the decision of
// v + 1 does not wait for the vector update.  A row is read from LDS only
// when its variable flips (most steps of an anneal reject).
#include "sa_real_block.h"

namespace misa {

namespace {

// A condition that holds in every lane or in none, as a scalar (so that the
// branch on it is a scalar branch).
__device__ __forceinline__ bool uniform(bool c) { return __builtin_amdgcn_ballot_w64(c) != 0ull; }

__device__ __forceinline__ double readlane_f64(double x, int l) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

// p[i] through a vector load made uniform: a pending scalar load would make
// every LDS wait in the sweep loop wait for all LDS operations.
__device__ __forceinline__ double uniform_load_f64(const double* p, int i) {
    asm("" : "+v"(i));
    const uint64_t b = __builtin_bit_cast(uint64_t, p[i]);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

}  // namespace

template <int NR, int PROP, bool COUNT>
__global__ __launch_bounds__(1024) void sa_dense_kernel(SaArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int ST = 64 * NR;  // row stride of the dense matrix (doubles)
    double* Jl = reinterpret_cast<double*>(smem);
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));

    // ---- stage the dense couplings --------------------------------------------
    {
        const double2* src = reinterpret_cast<const double2*>(a.dense);
        double2* dst = reinterpret_cast<double2*>(Jl);
        const int tot = n * (ST / 2);
        for (int i = threadIdx.x; i < tot; i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    const int rlocal = blockIdx.x * a.waves_per_block + wave;
    if (rlocal >= a.R) return;
    const int r = a.r0 + rlocal;

    // ---- spins in: bit l of M[i] set when variable 64 i + l is -1 -------------
    uint64_t M[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int v = 64 * i + lane;
        const int s = v < n ? (int)a.states[(size_t)r * n + v] : 1;
        M[i] = __builtin_amdgcn_ballot_w64(s < 0);
    }
    auto spin_of = [&](int u) -> double {
        const uint64_t m = NR == 1 || u < 64 ? M[0] : M[NR - 1];
        return (m >> (u & 63)) & 1ull ? -1.0 : 1.0;
    };
    // ---- fields, summed in adjacency order (cpu_sa.cpp:62-68) -------------------
    double f[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int v = 64 * i + lane;
        double e = 0.0;
        if (v < n) {
            e = a.h[v];
            for (int k = a.rowptr[v]; k < a.rowptr[v + 1]; ++k) e += spin_of(a.col[k]) * a.Jd[k];
        }
        f[i] = e;
    }

    uint64_t s0 = a.seeds[r];
    if (s0 == 0) s0 = ~0ull;
    uint64_t s1 = 0;
    // flips of this lane's variables (counters only)
    unsigned cnt[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) cnt[i] = 0;

    // carried from step to step: the field of the variable about to be
    // decided, already patched for its predecessor's flip
    double fC = readlane_f64(f[0], 0);

    for (int bi = 0; bi < a.nbetas; ++bi) {
        const double beta = uniform_load_f64(a.betas, bi);
        const double thr = uniform_load_f64(a.thr, bi);  // 44.36142 / beta, cpu_sa.cpp:131
        // cpu_sa.cpp:140-166: skip above the threshold, else Metropolis /
        // Gibbs (every value is the same in all lanes)
        auto decide = [&](const double fv, const double sv) -> bool {
            const double de = -2.0 * sv * fv;
            if (uniform(de >= thr)) return false;  // cpu_sa.cpp:140
            if constexpr (PROP == 1)
                if (uniform(de <= 0.0)) return true;
            const uint64_t x = real_detail::xs_step_r(s0, s1);
            const float xh = (float)(uint32_t)(x >> 40);
            if constexpr (PROP == 1) {
                const double t = -de * beta;
                const float a = __builtin_amdgcn_exp2f((float)t * 1.44269504f) * 16777216.0f;
                if (uniform(a * 0.99993896f > xh + 1.0f)) return true;
                if (uniform(a * 1.00006104f < xh)) return false;
                return uniform(exp(t) * real_detail::kTwo64 > (double)x);  // :154
            } else {
                const double t = de * beta;
                const float a =
                    16777216.0f * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f((float)t * 1.44269504f));
                if (uniform(a * 0.99993896f > xh + 1.0f)) return true;
                if (uniform(a * 1.00006104f < xh)) return false;
                return uniform(real_detail::kTwo64 > (double)x * (1.0 + exp(t)));  // :163
            }
        };
        // step of variable v (in register HV of lane v & 63) followed by nx
        // (register HN)
        auto step = [&](auto hv, auto hn, const int v, const int nx) {
            constexpr int HV = decltype(hv)::value, HN = decltype(hn)::value;
            double fN = readlane_f64(f[HN], nx & 63);
            const double sv = (M[HV] >> (v & 63)) & 1ull ? -1.0 : 1.0;
            if (decide(fC, sv)) {
                double Jv[NR];
#pragma unroll
                for (int i = 0; i < NR; ++i) Jv[i] = Jl[v * ST + 64 * i + lane];
                const double m2s = -2.0 * sv;  // field change per unit J: -2 s_v
#pragma unroll
                for (int i = 0; i < NR; ++i) f[i] = __builtin_fma(m2s, Jv[i], f[i]);
                fN = __builtin_fma(m2s, readlane_f64(Jv[HN], nx & 63), fN);  // J(v, nx)
                M[HV] ^= 1ull << (v & 63);
                if constexpr (COUNT) cnt[HV] += lane == (v & 63) ? 1u : 0u;
            }
            fC = fN;
        };
        using Z = std::integral_constant<int, 0>;
        using O = std::integral_constant<int, NR - 1>;
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            if constexpr (NR == 1) {
                for (int v = 0; v + 1 < n; ++v) step(Z{}, Z{}, v, v + 1);
                step(Z{}, Z{}, n - 1, 0);
            } else {
                for (int v = 0; v < 63; ++v) step(Z{}, Z{}, v, v + 1);
                step(Z{}, O{}, 63, 64);
                for (int v = 64; v + 1 < n; ++v) step(O{}, O{}, v, v + 1);
                step(O{}, Z{}, n - 1, 0);
            }
        }
    }

    // ---- spins out and energy (cpu_sa.cpp:210-227) ----------------------------
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        const int v = 64 * i + lane;
        if (v < n) a.states[(size_t)r * n + v] = (int8_t)((M[i] >> lane) & 1ull ? -1 : 1);
    }
    if (COUNT) {
        unsigned long long fl = 0, dg = 0;
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int v = 64 * i + lane;
            if (v < n) {
                fl += cnt[i];
                dg += (unsigned long long)cnt[i] * (unsigned long long)(a.rowptr[v + 1] - a.rowptr[v]);
            }
        }
        for (int o = 32; o >= 1; o >>= 1) {
            fl += __shfl_xor(fl, o);
            dg += __shfl_xor(dg, o);
        }
        if (lane == 0 && a.counters) {
            a.counters[2 * (size_t)r] = fl;
            a.counters[2 * (size_t)r + 1] = dg;
        }
    }
    if (lane == 0) {
        double en = 0.0;
        for (int v = 0; v < n; ++v) en += spin_of(v) * a.h[v];
        for (int64_t c = 0; c < a.m; ++c) en += spin_of(a.cs[c]) * a.cw[c] * spin_of(a.ce[c]);
        a.energies[r] = en;
    }
}

template <int NR, int PROP, bool COUNT>
static hipError_t launch_dense_variant(const SaArgs& a, int nblocks, hipStream_t stream) {
    auto fn = sa_dense_kernel<NR, PROP, COUNT>;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(64 * a.waves_per_block), a.lds_bytes, stream, a);
    return hipGetLastError();
}

template <int NR>
static hipError_t launch_dense_nr(const SaArgs& a, int proposal, int nblocks, hipStream_t stream) {
    if (proposal == 1)
        return a.counters ? launch_dense_variant<NR, 1, true>(a, nblocks, stream)
                          : launch_dense_variant<NR, 1, false>(a, nblocks, stream);
    return a.counters ? launch_dense_variant<NR, 0, true>(a, nblocks, stream)
                      : launch_dense_variant<NR, 0, false>(a, nblocks, stream);
}

hipError_t launch_sa_dense(const SaArgs& a, int proposal, int nblocks, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    if (a.dense_nr == 1) return launch_dense_nr<1>(a, proposal, nblocks, stream);
    if (a.dense_nr == 2) return launch_dense_nr<2>(a, proposal, nblocks, stream);
    return hipErrorInvalidValue;
}

}  // namespace misa
