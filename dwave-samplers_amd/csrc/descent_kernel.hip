// Steepest descent (greedy local search) for gfx950: one wave per read.
//
// MI355X counterpart of the reference's dense steepest-descent solver
// (dwave/samplers/greedy/src/descent.cpp:102-168): flip energies
// dE_v = (-2 s_v)(h_v + sum_u s_u J_vu) for the read's state, then repeatedly
// flip the variable with the most negative dE (lowest index on ties), negate
// its dE and subtract 4 s_best J s_u from each neighbour's dE, until no dE is
// negative.  The wave keeps dE (float64) and the spins in LDS, plus the
// minimum (value, index) of every 64-variable chunk; a step reduces the chunk
// minima, flips, updates the neighbours (one lane per adjacency entry, or one
// lane in adjacency order for rows that repeat a neighbour) and recomputes
// only the chunks it touched.  Self-loops are constant energy and are skipped.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "descent_engine.h"

namespace misa {
namespace {

struct VI {
    double v;
    int i;
};

// Lexicographic (value, index) minimum; NaN values never win.
__device__ __forceinline__ VI vi_min(VI a, VI b) {
    const bool take = (b.v < a.v) || (b.v == a.v && b.i < a.i) || (a.v != a.v);
    return take ? b : a;
}

__device__ __forceinline__ VI wave_min(VI x) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        VI o;
        o.v = __shfl_xor(x.v, off, 64);
        o.i = __shfl_xor(x.i, off, 64);
        x = vi_min(x, o);
    }
    return x;
}

__device__ __forceinline__ void chunk_update(const double* dE, double* cmin, int* cidx, int c, int n, int lane) {
    const int v = 64 * c + lane;
    VI x;
    x.v = v < n ? dE[v] : __builtin_inf();
    x.i = v < n ? v : 0x7fffffff;
    x = wave_min(x);
    if (lane == 0) {
        cmin[c] = x.v;
        cidx[c] = x.i;
    }
}

__global__ __launch_bounds__(256) void descent_kernel(DescentArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * a.waves + wave;
    if (r >= a.R) return;  // whole wave exits together (no block barrier below)
    const int n = a.n, nc = (n + 63) >> 6;
    unsigned char* base = smem + (size_t)wave * a.wave_lds;
    double* dE = reinterpret_cast<double*>(base);
    double* cmin = reinterpret_cast<double*>(base + 8 * (size_t)n);
    int* cidx = reinterpret_cast<int*>(base + 8 * (size_t)n + 8 * (size_t)nc);
    int8_t* s = reinterpret_cast<int8_t*>(base + 8 * (size_t)n + 12 * (size_t)nc);
    int8_t* gs = a.states + (size_t)r * n;

    for (int v = lane; v < n; v += 64) s[v] = gs[v];
    __builtin_amdgcn_wave_barrier();
    // descent.cpp:34-52 (self entries skipped)
    for (int v = lane; v < n; v += 64) {
        const double delta = -2 * s[v];
        double contrib = a.h[v];
        for (int k = a.rowptr[v]; k < a.rowptr[v + 1]; ++k) {
            const int u = a.col[k];
            if (u != v) contrib += s[u] * a.J[k];
        }
        dE[v] = delta * contrib;
    }
    __builtin_amdgcn_wave_barrier();
    for (int c = 0; c < nc; ++c) chunk_update(dE, cmin, cidx, c, n, lane);
    __builtin_amdgcn_wave_barrier();

    uint32_t steps = 0;
    while (steps < a.max_steps) {
        VI b;
        b.v = __builtin_inf();
        b.i = 0x7fffffff;
        for (int c = lane; c < nc; c += 64) {
            VI x;
            x.v = cmin[c];
            x.i = cidx[c];
            b = vi_min(b, x);
        }
        b = wave_min(b);
        if (!(b.v < 0.0)) break;
        const int best = b.i;
        const int sb = -s[best];  // the new spin of best
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            s[best] = (int8_t)sb;
            dE[best] = -dE[best];
        }
        __builtin_amdgcn_wave_barrier();
        const int rb = a.rowptr[best], re = a.rowptr[best + 1];
        if (a.serial[best]) {
            if (lane == 0)
                for (int k = rb; k < re; ++k) {
                    const int u = a.col[k];
                    if (u != best) dE[u] -= 4 * sb * a.J[k] * s[u];
                }
        } else {
            for (int k = rb + lane; k < re; k += 64) {
                const int u = a.col[k];
                if (u != best) dE[u] -= 4 * sb * a.J[k] * s[u];
            }
        }
        __builtin_amdgcn_wave_barrier();
        // Recompute every chunk the step touched, each once.
        for (int k0 = rb - 1; k0 < re; k0 += 64) {
            const int k = k0 + lane;
            int ch = -1;
            if (k == rb - 1 && lane == 0) ch = best >> 6;
            else if (k >= rb && k < re) ch = a.col[k] >> 6;
            uint64_t pending = __ballot(ch >= 0);
            while (pending) {
                const int src = __builtin_ctzll(pending);
                const int c = __shfl(ch, src, 64);
                chunk_update(dE, cmin, cidx, c, n, lane);
                pending &= __ballot(ch != c);
            }
        }
        __builtin_amdgcn_wave_barrier();
        ++steps;
    }
    __builtin_amdgcn_wave_barrier();
    for (int v = lane; v < n; v += 64) gs[v] = s[v];
    if (lane == 0) {
        // descent.cpp:65-85
        double e = 0.0;
        for (int v = 0; v < n; ++v) e += s[v] * a.h[v];
        for (int64_t c = 0; c < a.m; ++c) e += s[a.cs[c]] * a.cw[c] * s[a.ce[c]];
        a.energies[r] = e;
        a.steps[r] = steps;
    }
}

}  // namespace

hipError_t launch_descent(const DescentArgs& a, hipStream_t stream) {
    if (a.R <= 0) return hipSuccess;
    const int lds = a.waves * a.wave_lds;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(descent_kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    const int nblocks = (a.R + a.waves - 1) / a.waves;
    hipLaunchKernelGGL(descent_kernel, dim3(nblocks), dim3(64 * a.waves), lds, stream, a);
    return hipGetLastError();
}

}  // namespace misa
