// Discrete-imaginary-time path-integral Monte Carlo sweep kernel for gfx950.
//
// MI355X counterpart of the reference's single-qubit worldline update
// (dwave/samplers/sqa/src/localPIMC.cpp:350-383, driven by run() :128-160):
// a site's whole worldline is resampled from its conditional law.  Here the
// worldline has P <= 32 slices held as the bits of one word, so
//   * kinks (forced cuts) are w ^ rotl(w, 1);
//   * proposed cuts of equal-spin bonds (probability tanh(Hd/P) each, the
//     discrete form of the reference's Poisson break proposals :527-538) come
//     from geometric gaps;
//   * a segment's field, the sum over its slices of h_i + sum_j J_ij s_jk, is
//     a popcount of each neighbour word under the segment mask;
//   * the segment spin is drawn by heat bath, P(s) ~ exp(-(Hp/P) F s)
//     (the reference's GibbsSamplePM1, :551-556).
// Sites of one colour class are never adjacent; a block updates a whole class
// in parallel (lanes = (site, read) pairs) and synchronises between classes.
// Each (read, sweep, site) update draws from its own xoshiro128+ stream
// hashed from the read seed, so results do not depend on the launch shape.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pimc_engine.h"

namespace misa {
namespace {

__device__ __forceinline__ uint64_t pm_mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebull;
    z ^= z >> 31;
    return z;
}

__device__ __forceinline__ uint32_t pm_fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__device__ __forceinline__ uint32_t pm_rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

// xoshiro128+ (Blackman and Vigna): the sampler uses the upper 24 bits only.
struct Xs {
    uint32_t s0, s1, s2, s3;
    __device__ __forceinline__ uint32_t next() {
        const uint32_t r = s0 + s3;
        const uint32_t t = s1 << 9;
        s2 ^= s0;
        s3 ^= s1;
        s1 ^= s2;
        s0 ^= s3;
        s2 ^= t;
        s3 = pm_rotl(s3, 11);
        return r;
    }
};

__device__ __forceinline__ Xs task_rng(uint64_t key, uint32_t site) {
    Xs g;
    g.s0 = pm_fmix32((uint32_t)key + site);
    g.s1 = pm_fmix32((uint32_t)(key >> 32) ^ (0x9E3779B9u * (site + 1u)));
    g.s2 = g.s0 ^ 0x9E3779B9u;
    g.s3 = g.s1 ^ 0x7F4A7C15u;
    return g;
}

// 2^y, y in [-93, 93]: a degree-6 polynomial for 2^frac(y) and an exponent add.
__device__ __forceinline__ float pm_exp2f(float y) {
    const float fi = __builtin_floorf(y);
    const float f = y - fi;
    float p = 0x1.c6e292p-13f;
    p = __builtin_fmaf(p, f, 0x1.46301cp-10f);
    p = __builtin_fmaf(p, f, 0x1.3d24eap-7f);
    p = __builtin_fmaf(p, f, 0x1.c68562p-5f);
    p = __builtin_fmaf(p, f, 0x1.ebfd9ap-3f);
    p = __builtin_fmaf(p, f, 0x1.62e42ap-1f);
    p = __builtin_fmaf(p, f, 0x1p+0f);
    return __uint_as_float((uint32_t)((int32_t)__float_as_uint(p) + (int32_t)fi * (1 << 23)));
}

// log2(v * 2^-24) for v in [1, 2^24].
__device__ __forceinline__ float pm_log2u(uint32_t v) {
    const uint32_t bits = __float_as_uint((float)v);
    int e = (int)(bits >> 23) - 127;
    float mm = __uint_as_float((bits & 0x7fffffu) | 0x3f800000u);
    if (mm > 0x1.6a09e6p+0f) {
        mm = mm * 0.5f;
        e += 1;
    }
    const float t = mm - 1.0f;
    float q = -0x1.239648p-3f;
    q = __builtin_fmaf(q, t, 0x1.dc1d94p-3f);
    q = __builtin_fmaf(q, t, -0x1.fe9c4ap-3f);
    q = __builtin_fmaf(q, t, 0x1.26355p-2f);
    q = __builtin_fmaf(q, t, -0x1.70de64p-2f);
    q = __builtin_fmaf(q, t, 0x1.ec7536p-2f);
    q = __builtin_fmaf(q, t, -0x1.71552cp-1f);
    q = __builtin_fmaf(q, t, 0x1.715476p+0f);
    const float tq = t * q;
    return (float)(e - 24) + tq;
}

__device__ __forceinline__ int pm_gap(uint32_t u, float ginv) {
    float y = pm_log2u((u >> 8) + 1u) * ginv;
    y = __builtin_fminf(y, 64.0f);
    return (int)y;
}

constexpr int kCache = 8;  // neighbour words kept in registers (coupling table)

// One site update.  KIND: PIMC_TABLE (per-coupler weights, float64 field
// sums), PIMC_SIGNED (all |J| = J0: S = satisfied neighbour slices, counted
// per segment from LDS), PIMC_PLANES (all |J| = J0 and degree <= 15: the
// per-slice counts of satisfied neighbours are built once as 4 bit planes,
// and a segment's S is 4 masked popcounts).  Returns the slices changed.
template <int KIND>
__device__ __forceinline__ uint32_t site_update(const PimcArgs& a, uint32_t* __restrict__ sW,
                                                const int* __restrict__ srow, const uint16_t* __restrict__ scol,
                                                const uint16_t* __restrict__ sjx, const double* __restrict__ jt,
                                                const double* __restrict__ sh, int i, int rho, int lg, uint64_t key,
                                                double bs2l, float B, float ginv, int cut) {
    const int P = a.P;
    const uint32_t maskP = a.maskP;
    Xs g = task_rng(key, (uint32_t)i);
    const int wi = (i << lg) + rho;
    const uint32_t w0 = sW[wi];
    uint32_t w = w0;
    const int rb = srow[i];
    const int deg = srow[i + 1] - rb;
    const double hi = sh[i];
    float A = 0.0f;
    uint32_t p0 = 0u, p1 = 0u, p2 = 0u, p3 = 0u;
    uint32_t nw[kCache];
    double nj[kCache];
    if (KIND != PIMC_TABLE) A = (float)((hi - a.J0 * (double)deg) * bs2l);
    if (KIND == PIMC_PLANES) {
        for (int e = 0; e < deg; ++e) {
            const uint32_t cv = scol[rb + e];
            const uint32_t x = sW[((cv & 0x7fffu) << lg) + rho] ^ ((cv & 0x8000u) ? maskP : 0u);
            const uint32_t t0 = p0 & x;
            p0 ^= x;
            const uint32_t t1 = p1 & t0;
            p1 ^= t0;
            const uint32_t t2 = p2 & t1;
            p2 ^= t1;
            p3 ^= t2;
        }
    }
    if (KIND == PIMC_TABLE) {
#pragma unroll
        for (int e = 0; e < kCache; ++e) {
            nw[e] = 0u;
            nj[e] = 0.0;
            if (e < deg) {
                const uint32_t cv = scol[rb + e];
                nw[e] = sW[(cv << lg) + rho];
                nj[e] = jt[sjx[rb + e]];
            }
        }
    }

    uint32_t cuts = w ^ (((w << 1) | (w >> (P - 1))) & maskP);
    if (cut == PIMC_CUT_ALL) {
        cuts = maskP;
    } else if (cut == PIMC_CUT_PROP) {
        int pos = pm_gap(g.next(), ginv);
        while (pos < P) {
            cuts |= 1u << pos;
            pos += 1 + pm_gap(g.next(), ginv);
        }
    }
    if (__builtin_popcount(cuts) <= 1) cuts = 1u;
    const int first = __builtin_ctz(cuts);
    uint32_t rem = cuts & (cuts - 1u);
    int s = first;
    for (;;) {
        int b;
        if (rem) {
            b = __builtin_ctz(rem);
            rem &= rem - 1u;
        } else {
            b = first + P;
        }
        const int len = b - s;
        const uint64_t m64 = (1ull << b) - (1ull << s);
        const uint32_t m = (uint32_t)((m64 | (m64 >> P)) & maskP);
        float y;
        if (KIND == PIMC_PLANES) {
            const int S = __builtin_popcount(p0 & m) + 2 * __builtin_popcount(p1 & m) +
                          4 * __builtin_popcount(p2 & m) + 8 * __builtin_popcount(p3 & m);
            const float t = (float)S * B;
            y = __builtin_fmaf((float)len, A, t);
        } else if (KIND == PIMC_SIGNED) {
            int S = 0;
            for (int e = 0; e < deg; ++e) {
                const uint32_t cv = scol[rb + e];
                S += __builtin_popcount((sW[((cv & 0x7fffu) << lg) + rho] ^ ((cv & 0x8000u) ? maskP : 0u)) & m);
            }
            const float t = (float)S * B;
            y = __builtin_fmaf((float)len, A, t);
        } else {
            double F = (double)len * hi;
#pragma unroll
            for (int e = 0; e < kCache; ++e)
                if (e < deg) F = F + nj[e] * (double)(2 * __builtin_popcount(nw[e] & m) - len);
            for (int e = kCache; e < deg; ++e) {
                const uint32_t cv = scol[rb + e];
                F = F + jt[sjx[rb + e]] * (double)(2 * __builtin_popcount(sW[(cv << lg) + rho] & m) - len);
            }
            y = (float)(F * bs2l);
        }
        y = __builtin_fminf(__builtin_fmaxf(y, -92.0f), 92.0f);
        const float ex = pm_exp2f(y);
        const float u = __uint_as_float(0x3F800000u | (g.next() >> 9)) - 1.0f;
        const float t = 1.0f + ex;
        const bool plus = (u * t) < 1.0f;
        w = (w & ~m) | (plus ? m : 0u);
        if (b == first + P) break;
        s = b;
    }
    sW[wi] = w;
    return (uint32_t)__builtin_popcount(w ^ w0);
}

__device__ __forceinline__ int chain_find(uint16_t* parent, int x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

// Segment of a qubit with these cuts that holds slice s: the last cut at or
// below s, else the last segment (it wraps past slice P - 1).
__device__ __forceinline__ int chain_seg_of(uint32_t cuts, int s) {
    const uint32_t upto = s >= 31 ? 0xffffffffu : ((2u << s) - 1u);
    const int r = __builtin_popcount(cuts & upto) - 1;
    return r >= 0 ? r : __builtin_popcount(cuts) - 1;
}

// One chain update (the reference's chainUpdate, localPIMC.cpp:385-452, with
// makeDomainGraph :462-525): chain c is qubits c*Q .. c*Q+Q-1.  The segments
// of every qubit are the nodes; consecutive qubits' segments are joined
// piece by piece of their joint cuts where their spins agree, with
// probability 1 - exp(2 Jchain (Hp/P) len); each component takes one
// heat-bath spin from the sum of its segments' fields (neighbour worldlines as
// before the update, float64, added in node order; a component's root is its
// smallest node).  The random draws, in order: proposed cuts qubit by qubit,
// one per join test where the spins agree, one per component.
__device__ __noinline__ uint32_t chain_update(const PimcArgs& a, uint32_t* __restrict__ sW,
                                              const int* __restrict__ srow, const uint16_t* __restrict__ scol,
                                              const uint16_t* __restrict__ sjx, const double* __restrict__ jt,
                                              const double* __restrict__ sh, int c, int rho, int lg, uint64_t key,
                                              double bs2l, float ginv, int cut) {
    const int P = a.P, Q = a.Q, q0 = c * Q;
    const uint32_t maskP = a.maskP;
    uint32_t cuts[kChainQMax];
    uint16_t parent[kChainNodes];
    double fc[kChainNodes];
    uint32_t up[kChainNodes / 32];
    Xs g = task_rng(key, (uint32_t)c);
    int N = 0;
    for (int q = 0; q < Q; ++q) {
        const uint32_t w = sW[((q0 + q) << lg) + rho];
        uint32_t cu = w ^ (((w << 1) | (w >> (P - 1))) & maskP);
        if (cut == PIMC_CUT_ALL) {
            cu = maskP;
        } else if (cut == PIMC_CUT_PROP) {
            int pos = pm_gap(g.next(), ginv);
            while (pos < P) {
                cu |= 1u << pos;
                pos += 1 + pm_gap(g.next(), ginv);
            }
        }
        if (__builtin_popcount(cu) <= 1) cu = 1u;
        cuts[q] = cu;
        N += __builtin_popcount(cu);
    }
    for (int x = 0; x < N; ++x) parent[x] = (uint16_t)x;
    for (int x = 0; x < (N + 31) / 32; ++x) up[x] = 0u;
    // joins
    int bq = 0;
    for (int q = 0; q + 1 < Q; ++q) {
        const int bq1 = bq + __builtin_popcount(cuts[q]);
        const uint32_t U = cuts[q] | cuts[q + 1];
        const uint32_t diff = sW[((q0 + q) << lg) + rho] ^ sW[((q0 + q + 1) << lg) + rho];
        const int first = __builtin_ctz(U);
        uint32_t rem = U & (U - 1u);
        int s = first;
        for (;;) {
            int b;
            if (rem) {
                b = __builtin_ctz(rem);
                rem &= rem - 1u;
            } else {
                b = first + P;
            }
            if (((diff >> s) & 1u) == 0u) {
                float y = (float)(a.jchain * (double)(b - s) * bs2l);
                y = __builtin_fminf(__builtin_fmaxf(y, -92.0f), 92.0f);
                const float pnot = pm_exp2f(y);
                const float u = __uint_as_float(0x3F800000u | (g.next() >> 9)) - 1.0f;
                if (pnot < u) {
                    const int x = chain_find(parent, bq + chain_seg_of(cuts[q], s));
                    const int z = chain_find(parent, bq1 + chain_seg_of(cuts[q + 1], s));
                    if (x != z) parent[x > z ? x : z] = (uint16_t)(x < z ? x : z);
                }
            }
            if (b == first + P) break;
            s = b;
        }
        bq = bq1;
    }
    // node fields, summed per component in node order
    int x = 0;
    for (int q = 0; q < Q; ++q) {
        const int i = q0 + q;
        const uint32_t cu = cuts[q];
        const int rb = srow[i], re = srow[i + 1];
        const double hi = sh[i];
        const int first = __builtin_ctz(cu);
        uint32_t rem = cu & (cu - 1u);
        int s = first;
        for (;;) {
            int b;
            if (rem) {
                b = __builtin_ctz(rem);
                rem &= rem - 1u;
            } else {
                b = first + P;
            }
            const int len = b - s;
            const uint64_t m64 = (1ull << b) - (1ull << s);
            const uint32_t m = (uint32_t)((m64 | (m64 >> P)) & maskP);
            double F = (double)len * hi;
            for (int e = rb; e < re; ++e)
                F = F + jt[sjx[e]] * (double)(2 * __builtin_popcount(sW[((int)scol[e] << lg) + rho] & m) - len);
            const int r = chain_find(parent, x);
            fc[r] = r == x ? F : fc[r] + F;
            ++x;
            if (b == first + P) break;
            s = b;
        }
    }
    // one spin per component, in order of the roots
    for (x = 0; x < N; ++x) {
        if (parent[x] != x) continue;
        float y = (float)(fc[x] * bs2l);
        y = __builtin_fminf(__builtin_fmaxf(y, -92.0f), 92.0f);
        const float ex = pm_exp2f(y);
        const float u = __uint_as_float(0x3F800000u | (g.next() >> 9)) - 1.0f;
        if ((u * (1.0f + ex)) < 1.0f) up[x >> 5] |= 1u << (x & 31);
    }
    // new worldlines
    uint32_t changed = 0;
    x = 0;
    for (int q = 0; q < Q; ++q) {
        const uint32_t cu = cuts[q];
        const int first = __builtin_ctz(cu);
        uint32_t rem = cu & (cu - 1u);
        int s = first;
        uint32_t w = 0;
        for (;;) {
            int b;
            if (rem) {
                b = __builtin_ctz(rem);
                rem &= rem - 1u;
            } else {
                b = first + P;
            }
            const uint64_t m64 = (1ull << b) - (1ull << s);
            const uint32_t m = (uint32_t)((m64 | (m64 >> P)) & maskP);
            const int r = chain_find(parent, x);
            if ((up[r >> 5] >> (r & 31)) & 1u) w |= m;
            ++x;
            if (b == first + P) break;
            s = b;
        }
        const int wi = ((q0 + q) << lg) + rho;
        changed += (uint32_t)__builtin_popcount(w ^ sW[wi]);
        sW[wi] = w;
    }
    return changed;
}

template <int KIND>
__global__ __launch_bounds__(1024) void pimc_kernel(PimcArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = a.n, lg = a.rpb_log2, rpb = 1 << lg;
    const int rb0 = blockIdx.x << lg;
    const int nr = min(rpb, a.R - rb0);
    double* sh = reinterpret_cast<double*>(smem + a.off_h);
    double* sjt = reinterpret_cast<double*>(smem + a.off_jtab);
    uint32_t* sW = reinterpret_cast<uint32_t*>(smem + a.off_W);
    int* srow = reinterpret_cast<int*>(smem + a.off_rowptr);
    int* scoff = reinterpret_cast<int*>(smem + a.off_coff);
    uint16_t* scol = reinterpret_cast<uint16_t*>(smem + a.off_col);
    uint16_t* sjx = reinterpret_cast<uint16_t*>(smem + a.off_jidx);
    uint16_t* ssite = reinterpret_cast<uint16_t*>(smem + a.off_sites);

    for (int t = tid; t < n; t += nt) {
        sh[t] = a.h[t];
        ssite[t] = a.class_sites[t];
    }
    for (int t = tid; t <= n; t += nt) srow[t] = a.rowptr[t];
    for (int t = tid; t <= a.nclass; t += nt) scoff[t] = a.class_off[t];
    for (int t = tid; t < a.m2; t += nt) scol[t] = a.col[t];
    if (KIND == PIMC_TABLE || KIND == PIMC_CHAIN) {
        for (int t = tid; t < a.m2; t += nt) sjx[t] = a.jidx[t];
        if (a.jtab_in_lds)
            for (int t = tid; t < a.ntab; t += nt) sjt[t] = a.jtab[t];
    }
    for (int t = tid; t < (n << lg); t += nt) {
        const int r = t / n, i = t - r * n;
        sW[(i << lg) + r] = r < nr ? a.W[(size_t)(rb0 + r) * n + i] : 0u;
    }
    __syncthreads();
    const double* jt = a.jtab_in_lds ? sjt : a.jtab;

    const int rho = tid & (rpb - 1);
    const bool own = rho < nr;
    const uint64_t seed = own ? a.seeds[rb0 + rho] : 0ull;
    uint64_t nflip = 0;
    for (int k = 0; k < a.nsteps; ++k) {
        const double bs2l = 2.0 * a.bs[k] * 0x1.71547652b82fep+0;  // 2 (Hp/P) log2(e)
        const float B = (float)(2.0 * a.J0 * bs2l);
        const float ginv = a.ginv[k];
        const int cut = a.cut[k];
        // Chain updates: Q passes over the chains per sweep (about one update
        // per qubit, as localPIMC.cpp:146-152 makes n per sweep).
        const int npass = KIND == PIMC_CHAIN ? a.Q : 1;
        for (int sw = 0; sw < a.spb; ++sw)
            for (int pass = 0; pass < npass; ++pass) {
                const uint64_t gidx = ((uint64_t)k * (uint64_t)a.spb + (uint64_t)sw) * (uint64_t)npass + pass;
                const uint64_t key = pm_mix64(seed + (gidx + 1ull) * 0x9E3779B97F4A7C15ull);
                for (int c = 0; c < a.nclass; ++c) {
                    const int lo = scoff[c];
                    const int ntask = (scoff[c + 1] - lo) << lg;
                    if (own) {
                        uint32_t f = 0;
                        for (int t = tid; t < ntask; t += nt) {
                            if constexpr (KIND == PIMC_CHAIN)
                                f += chain_update(a, sW, srow, scol, sjx, jt, sh, ssite[lo + (t >> lg)], rho, lg,
                                                  key, bs2l, ginv, cut);
                            else
                                f += site_update<KIND>(a, sW, srow, scol, sjx, jt, sh, ssite[lo + (t >> lg)], rho,
                                                       lg, key, bs2l, B, ginv, cut);
                        }
                        nflip += f;
                    }
                    __syncthreads();
                }
            }
        if (a.stats && a.interval > 0 && k % a.interval == 0) {
            const int colk = k / a.interval;
            for (int t = tid; t < (n << lg); t += nt) {
                const int r = t / n, i = t - r * n;
                if (r < nr)
                    a.stats[((size_t)(rb0 + r) * a.ncol + colk) * n + i] = (sW[(i << lg) + r] & 1u) ? 1 : -1;
            }
            __syncthreads();
        }
    }
    for (int t = tid; t < (n << lg); t += nt) {
        const int r = t / n, i = t - r * n;
        if (r < nr) a.W[(size_t)(rb0 + r) * n + i] = sW[(i << lg) + r];
    }
    if (a.flips && own && nflip) atomicAdd(reinterpret_cast<unsigned long long*>(a.flips + rb0 + rho),
                                           (unsigned long long)nflip);
}

__global__ __launch_bounds__(64) void pimc_project_kernel(const uint32_t* __restrict__ W, int R, int n,
                                                          const double* __restrict__ h, int64_t m,
                                                          const int* __restrict__ cs, const int* __restrict__ ce,
                                                          const double* __restrict__ cw, int8_t* states,
                                                          double* energies) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint32_t* w = W + (size_t)r * n;
    if (states)
        for (int i = 0; i < n; ++i) states[(size_t)r * n + i] = (w[i] & 1u) ? 1 : -1;
    if (energies) {
        double e = 0.0;
        for (int v = 0; v < n; ++v) e += (double)((w[v] & 1u) ? 1 : -1) * h[v];
        for (int64_t c = 0; c < m; ++c)
            e += (double)((w[cs[c]] & 1u) ? 1 : -1) * cw[c] * (double)((w[ce[c]] & 1u) ? 1 : -1);
        energies[r] = e;
    }
}

}  // namespace

hipError_t launch_pimc(const PimcArgs& a, int kind, hipStream_t stream) {
    if (a.R <= 0) return hipSuccess;
    auto fn = kind == PIMC_PLANES   ? pimc_kernel<PIMC_PLANES>
              : kind == PIMC_SIGNED ? pimc_kernel<PIMC_SIGNED>
              : kind == PIMC_CHAIN  ? pimc_kernel<PIMC_CHAIN>
                                    : pimc_kernel<PIMC_TABLE>;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    const int nblocks = (a.R + (1 << a.rpb_log2) - 1) >> a.rpb_log2;
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(a.threads), a.lds_bytes, stream, a);
    return hipGetLastError();
}

hipError_t launch_pimc_project(const uint32_t* W, int R, int n, const double* h, int64_t m, const int* cs,
                               const int* ce, const double* cw, int8_t* states, double* energies,
                               hipStream_t stream) {
    if (R <= 0) return hipSuccess;
    hipLaunchKernelGGL(pimc_project_kernel, dim3((R + 63) / 64), dim3(64), 0, stream, W, R, n, h, m, cs, ce, cw,
                       states, energies);
    return hipGetLastError();
}

}  // namespace misa
