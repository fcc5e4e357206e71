// Simulated-annealing sweep kernels for gfx950 (CDNA4).
//
// Reference algorithm: dwave/samplers/sa/src/cpu_sa.cpp
//   get_flip_energy            :54-74   -> field initialisation below
//   simulated_annealing_run    :93-197  -> the beta / sweep / variable loops
//   get_state_energy           :210-227 -> epilogue energy
//   xorshift128+ FASTRAND      :27-34, seeding :283-284
//
// Layout (one wave = one block; 64 lanes split into groups of L = 2^lanes_log2
// lanes, one group per read, rpb = 64 / L reads per block):
//   * Every read keeps the reference's serial single-spin order and its own
//     xorshift128+ stream; the L lanes of a group replay the same decision
//     (so their RNG states stay identical) and split the neighbour updates
//     of an accepted flip across lanes (one CSR entry per lane).
//   * Per-read state is kept as the local field f_v = h_v + sum_k J_k s_k
//     instead of the reference's dE_v = -2 s_v f_v.  Scaling by -2 s_v is
//     exact in binary floating point, so every dE the reference forms is
//     reproduced bit for bit, and an accepted flip of v changes each
//     neighbour's field by -2 J s_v without reading the neighbour's spin.
//       - integer modes (all biases integers, |f| bounded): one packed word
//         p = 2 f + b per variable (b = 1 for spin -1), int8 or int16.
//         Acceptance uses a host-built table T[beta][dE/2] with "accept iff
//         rand < T", equal to the reference's double comparison for every
//         64-bit rand (no device exp).
//       - real mode: f64 field plus an int8 spin; acceptance evaluates the
//         reference's double expressions with the device exp.
//     Storage is variable-major, read-minor: word (v, g) at v * rpb + g.
//   * The state lives in LDS when rpb reads fit, else in device memory.
//   * Adjacency (CSR in coupler insertion order, cpu_sa.cpp:298-316) is read
//     from device memory; with sequential order the row bounds are
//     wave-uniform and load through the scalar cache.
//   * Rows holding a repeated neighbour or the variable itself (self-loops)
//     are updated by one lane in CSR order, which keeps the reference's
//     floating-point order for repeated entries.
#include <climits>
#include <type_traits>

#include "sa_engine.h"

namespace misa {

__device__ __forceinline__ uint64_t xs_step(uint64_t& s0, uint64_t& s1) {
    // xorshift128+, cpu_sa.cpp:27-34.
    uint64_t a = s0;
    const uint64_t b = s1;
    s0 = b;
    a ^= a << 23;
    s1 = a ^ b ^ (a >> 17) ^ (b >> 26);
    return s1 + b;
}

// 2^64 as a double: the value (double)RANDMAX takes in cpu_sa.cpp:154,163.
#define MISA_TWO64 18446744073709551616.0

template <typename ST, int ORDER, int PROP, bool LDS>
__global__ __launch_bounds__(64) void sa_sweep_kernel(SaArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool REAL = std::is_same<ST, double>::value;

    const int tid = threadIdx.x;
    const int lg = a.lanes_log2;
    const int L = 1 << lg;
    const int g = tid >> lg;
    const int j = tid & (L - 1);
    const int rpb = a.rpb;
    const int n = a.n;
    const int rlocal = blockIdx.x * rpb + g;
    const bool act = rlocal < a.R;
    const int r = a.r0 + rlocal;

    ST* F;
    int8_t* S;
    if constexpr (LDS) {
        F = reinterpret_cast<ST*>(smem);
        S = reinterpret_cast<int8_t*>(smem + (size_t)n * rpb * sizeof(ST));
    } else {
        unsigned char* base = reinterpret_cast<unsigned char*>(a.gstate) +
                              (size_t)blockIdx.x * rpb * state_bytes_per_read(
                                  REAL ? MODE_REAL : (sizeof(ST) == 1 ? MODE_INT8 : MODE_INT16), n);
        F = reinterpret_cast<ST*>(base);
        S = reinterpret_cast<int8_t*>(base + (size_t)n * rpb * sizeof(ST));
    }

    // ---- spins in ----------------------------------------------------------
    for (int v = j; v < n; v += L) {
        const int s = act ? (int)a.states[(size_t)r * n + v] : 1;
        if constexpr (REAL) S[v * rpb + g] = (int8_t)s;
        else F[v * rpb + g] = (ST)(s < 0 ? 1 : 0);
    }
    __syncthreads();
    // ---- fields, summed in adjacency order (cpu_sa.cpp:62-68) --------------
    for (int v = j; v < n; v += L) {
        const int kb = a.rowptr[v], ke = a.rowptr[v + 1];
        if constexpr (REAL) {
            double e = a.h[v];
            for (int k = kb; k < ke; ++k) e += (double)S[a.col[k] * rpb + g] * a.Jd[k];
            F[v * rpb + g] = e;
        } else {
            int e = (int)a.h[v];
            for (int k = kb; k < ke; ++k) {
                const int su = 1 - 2 * ((int)F[a.col[k] * rpb + g] & 1);
                e += su * (int)a.Ji[k];
            }
            const int b = (int)F[v * rpb + g] & 1;
            F[v * rpb + g] = (ST)(2 * e + b);
        }
    }
    __syncthreads();
    if (!act) return;

    // ---- anneal (cpu_sa.cpp:120-193) --------------------------------------
    uint64_t s0 = a.seeds[r];
    if (s0 == 0) s0 = ~0ull;
    uint64_t s1 = 0;
    unsigned long long nflip = 0, ndeg = 0;
    const int twidth = 2 * a.fmax + 1;

    for (int bi = 0; bi < a.nbetas; ++bi) {
        const double beta = a.betas[bi];
        double thr = 0.0;
        int kthr = 0;
        const uint64_t* trow = nullptr;
        if constexpr (REAL) {
            thr = a.thr[bi];
        } else {
            kthr = a.kthr[bi];
            trow = a.ttab + (size_t)bi * twidth + a.fmax;
        }
        for (int sw = 0; sw < a.sweeps_per_beta; ++sw) {
            for (int it = 0; it < n; ++it) {
                int v;
                if constexpr (ORDER == 1) {
                    const uint64_t x = xs_step(s0, s1);
                    v = (int)(x % (uint64_t)n);
                } else {
                    v = it;
                }
                bool flip = false;
                int sv;
                int pv = 0;
                if constexpr (REAL) {
                    const double f = F[v * rpb + g];
                    sv = S[v * rpb + g];
                    const double de = (double)(-2 * sv) * f;
                    if (!(de >= thr)) {
                        if constexpr (PROP == 1) {
                            if (de <= 0.0) {
                                flip = true;
                            } else {
                                const uint64_t x = xs_step(s0, s1);
                                flip = exp(-de * beta) * MISA_TWO64 > (double)x;
                            }
                        } else {
                            const uint64_t x = xs_step(s0, s1);
                            flip = MISA_TWO64 > (double)x * (1.0 + exp(de * beta));
                        }
                    }
                } else {
                    pv = (int)F[v * rpb + g];
                    const int fb = pv >> 1;
                    const int bb = pv & 1;
                    sv = 1 - 2 * bb;
                    const int k = bb ? fb : -fb;  // dE / 2
                    if (k < kthr) {
                        if constexpr (PROP == 1) {
                            if (k <= 0) {
                                flip = true;
                            } else {
                                const uint64_t x = xs_step(s0, s1);
                                flip = x < trow[k];
                            }
                        } else {
                            const uint64_t x = xs_step(s0, s1);
                            flip = x < trow[k];
                        }
                    }
                }
                if (flip) {
                    const int kb = a.rowptr[v], ke = a.rowptr[v + 1];
                    nflip += 1;
                    ndeg += (unsigned long long)(ke - kb);
                    if (a.serial_row[v]) {
                        if (j == 0) {
                            for (int k = kb; k < ke; ++k) {
                                const int u = a.col[k];
                                if constexpr (REAL) F[u * rpb + g] += (double)(-2 * sv) * a.Jd[k];
                                else F[u * rpb + g] = (ST)((int)F[u * rpb + g] - 4 * sv * (int)a.Ji[k]);
                            }
                            if constexpr (REAL) S[v * rpb + g] = (int8_t)(-sv);
                            else F[v * rpb + g] = (ST)((int)F[v * rpb + g] ^ 1);
                        }
                    } else {
                        for (int k = kb + j; k < ke; k += L) {
                            const int u = a.col[k];
                            if constexpr (REAL) F[u * rpb + g] += (double)(-2 * sv) * a.Jd[k];
                            else F[u * rpb + g] = (ST)((int)F[u * rpb + g] - 4 * sv * (int)a.Ji[k]);
                        }
                        if constexpr (REAL) S[v * rpb + g] = (int8_t)(-sv);
                        else F[v * rpb + g] = (ST)(pv ^ 1);
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
    }

    // ---- spins out and energy (cpu_sa.cpp:210-227) -------------------------
    auto spin = [&](int v) -> int {
        if constexpr (REAL) return (int)S[v * rpb + g];
        else return 1 - 2 * ((int)F[v * rpb + g] & 1);
    };
    for (int v = j; v < n; v += L) a.states[(size_t)r * n + v] = (int8_t)spin(v);
    if (j == 0) {
        double e = 0.0;
        for (int v = 0; v < n; ++v) e += (double)spin(v) * a.h[v];
        for (int64_t c = 0; c < a.m; ++c)
            e += (double)spin(a.cs[c]) * a.cw[c] * (double)spin(a.ce[c]);
        a.energies[r] = e;
        if (a.counters) {
            a.counters[2 * (size_t)r] = nflip;
            a.counters[2 * (size_t)r + 1] = ndeg;
        }
    }
}

template <typename ST, int ORDER, int PROP, bool LDS>
static hipError_t launch_variant(const SaArgs& a, int nblocks, hipStream_t stream) {
    auto fn = sa_sweep_kernel<ST, ORDER, PROP, LDS>;
    size_t lds = 0;
    if (LDS) {
        const int mode = std::is_same<ST, double>::value ? MODE_REAL
                                                          : (sizeof(ST) == 1 ? MODE_INT8 : MODE_INT16);
        lds = (size_t)a.rpb * state_bytes_per_read(mode, a.n);
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(64), lds, stream, a);
    return hipGetLastError();
}

template <typename ST, int ORDER, int PROP>
static hipError_t launch_lds(const SaArgs& a, bool lds, int nblocks, hipStream_t s) {
    return lds ? launch_variant<ST, ORDER, PROP, true>(a, nblocks, s)
               : launch_variant<ST, ORDER, PROP, false>(a, nblocks, s);
}

template <typename ST>
static hipError_t launch_mode(const SaArgs& a, int order, int prop, bool lds, int nb, hipStream_t s) {
    if (order == 0) {
        return prop == 1 ? launch_lds<ST, 0, 1>(a, lds, nb, s) : launch_lds<ST, 0, 0>(a, lds, nb, s);
    }
    return prop == 1 ? launch_lds<ST, 1, 1>(a, lds, nb, s) : launch_lds<ST, 1, 0>(a, lds, nb, s);
}

hipError_t launch_sa(const SaArgs& a, int mode, int order, int proposal, bool lds_state,
                     int nblocks, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    switch (mode) {
        case MODE_INT8: return launch_mode<int8_t>(a, order, proposal, lds_state, nblocks, stream);
        case MODE_INT16: return launch_mode<int16_t>(a, order, proposal, lds_state, nblocks, stream);
        default: return launch_mode<double>(a, order, proposal, lds_state, nblocks, stream);
    }
}

size_t max_dynamic_lds() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 65536;
    return (size_t)v;
}

}  // namespace misa
