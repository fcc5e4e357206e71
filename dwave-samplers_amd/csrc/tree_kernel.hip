// Exact inference by bucket elimination for gfx950: forward (log-sum-product)
// messages, downward messages, variable and pair marginals, and exact
// Boltzmann samples by backward sampling.  Counterpart of the reference's
// tree decomposition sampler (dwave/samplers/tree/samplers.py:186-359, its
// bucket-tree engine in tree/src/include: sample.hpp, buckettree.h,
// operations/logsumprod.h).  Table layout: see tree_engine.h.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tree_engine.h"

namespace misa {
namespace {

constexpr double kNegInf = -__builtin_inf();

// Online log-sum-exp accumulator (plain data: it also lives in LDS).
struct Lse {
    double m, s;
    __device__ void add(double v) {
        if (v > m) {
            s = s * exp(m - v) + 1.0;
            m = v;
        } else if (v > kNegInf) {
            s += exp(v - m);
        }
    }
    __device__ void merge(Lse o) {
        if (o.m > m) {
            s = s * exp(m - o.m) + o.s;
            m = o.m;
        } else if (o.m > kNegInf) {
            s += o.s * exp(o.m - m);
        }
    }
    __device__ double value() const { return m == kNegInf ? kNegInf : m + log(s); }
};

__device__ __forceinline__ Lse lse_empty() {
    Lse l;
    l.m = kNegInf;
    l.s = 0.0;
    return l;
}

__device__ __forceinline__ int64_t gather_bits(int64_t idx, const int* pos, int r) {
    int64_t j = 0;
    for (int t = 0; t < r; ++t) j |= ((idx >> pos[t]) & 1) << t;
    return j;
}

// T_b over the scope and mu_b = log-sum-exp over bit 0; one thread per
// separator index.
__global__ __launch_bounds__(256) void tree_forward_kernel(TreeArgs a, int b) {
    const TreeBucket B = a.buckets[b];
    const int64_t half = (int64_t)1 << (B.k - 1);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v[2];
#pragma unroll
        for (int xv = 0; xv < 2; ++xv) {
            const int64_t idx = (i << 1) | xv;
            double s = 0.0;
            for (int t = 0; t < B.nterms; ++t) {
                const TreeTerm tm = a.terms[B.term_off + t];
                const double xa = ((idx >> tm.pa) & 1) ? 1.0 : a.low;
                const double xb = tm.pb < 0 ? 1.0 : (((idx >> tm.pb) & 1) ? 1.0 : a.low);
                s += -a.beta * tm.coef * xa * xb;
            }
            for (int c = 0; c < B.nchild; ++c) {
                const TreeBucket C = a.buckets[a.children[B.child_off + c]];
                s += a.mu[C.mu_off + gather_bits(idx, a.pos + C.pos_off, C.k - 1)];
            }
            v[xv] = s;
            a.T[B.t_off + idx] = s;
        }
        const double m = v[0] > v[1] ? v[0] : v[1];
        a.mu[B.mu_off + i] = m + log(exp(v[0] - m) + exp(v[1] - m));
    }
}

// delta_c over child c's separator: the parent's belief T_p + delta_p summed
// (log-sum-exp) over the parent positions outside the separator, minus mu_c.
__global__ __launch_bounds__(256) void tree_downward_kernel(TreeArgs a, int c) {
    const TreeBucket C = a.buckets[c];
    const TreeBucket P = a.buckets[C.parent];
    const int r = C.k - 1;
    const int F = P.k - r;
    const int* pos = a.pos + C.pos_off;
    const int* fr = a.pos + C.free_off;
    const int64_t nj = (int64_t)1 << r;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nj; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t base = 0;
        for (int t = 0; t < r; ++t)
            if ((j >> t) & 1) base |= (int64_t)1 << pos[t];
        Lse acc = lse_empty();
        for (int64_t f = 0; f < ((int64_t)1 << F); ++f) {
            int64_t idx = base;
            for (int u = 0; u < F; ++u)
                if ((f >> u) & 1) idx |= (int64_t)1 << fr[u];
            acc.add(a.T[P.t_off + idx] + a.delta[P.d_off + (idx >> 1)]);
        }
        a.delta[C.d_off + j] = acc.value() - a.mu[C.mu_off + j];
    }
}

__device__ Lse block_lse(Lse x, Lse* sh) {
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x].merge(sh[threadIdx.x + w]);
        __syncthreads();
    }
    Lse r = sh[0];
    __syncthreads();
    return r;
}

// P(var_b high) from bucket b's belief, one block per bucket.
__global__ __launch_bounds__(256) void tree_var_marginal_kernel(TreeArgs a, double* out) {
    __shared__ Lse sh[256];
    const TreeBucket B = a.buckets[blockIdx.x];
    Lse l0 = lse_empty(), l1 = lse_empty();
    const int64_t size = (int64_t)1 << B.k;
    for (int64_t idx = threadIdx.x; idx < size; idx += blockDim.x) {
        const double v = a.T[B.t_off + idx] + a.delta[B.d_off + (idx >> 1)];
        if (idx & 1)
            l1.add(v);
        else
            l0.add(v);
    }
    l0 = block_lse(l0, sh);
    l1 = block_lse(l1, sh);
    if (threadIdx.x == 0) {
        const double x0 = l0.value(), x1 = l1.value();
        const double m = x0 > x1 ? x0 : x1;
        out[B.var] = exp(x1 - m) / (exp(x0 - m) + exp(x1 - m));
    }
}

// The four joint probabilities of (u, v) from the bucket holding their term.
__global__ __launch_bounds__(256) void tree_pair_marginal_kernel(TreeArgs a, const int* pairs, double* out) {
    __shared__ Lse sh[256];
    const int b = pairs[3 * blockIdx.x], pu = pairs[3 * blockIdx.x + 1], pv = pairs[3 * blockIdx.x + 2];
    const TreeBucket B = a.buckets[b];
    Lse l[4] = {lse_empty(), lse_empty(), lse_empty(), lse_empty()};
    const int64_t size = (int64_t)1 << B.k;
    for (int64_t idx = threadIdx.x; idx < size; idx += blockDim.x) {
        const double v = a.T[B.t_off + idx] + a.delta[B.d_off + (idx >> 1)];
        const int q = (int)((idx >> pu) & 1) + 2 * (int)((idx >> pv) & 1);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (c == q) l[c].add(v);
    }
    double x[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) x[c] = block_lse(l[c], sh).value();
    if (threadIdx.x == 0) {
        double m = x[0];
        for (int c = 1; c < 4; ++c) m = x[c] > m ? x[c] : m;
        double z = 0.0;
        for (int c = 0; c < 4; ++c) z += exp(x[c] - m);
        for (int c = 0; c < 4; ++c) out[4 * (size_t)blockIdx.x + c] = exp(x[c] - m) / z;
    }
}

__device__ __forceinline__ uint64_t xs_next(uint64_t& s0, uint64_t& s1) {
    uint64_t x = s0;
    const uint64_t y = s1;
    s0 = y;
    x ^= x << 23;
    s1 = x ^ y ^ (x >> 17) ^ (y >> 26);
    return s1 + y;
}

// splitmix64 output i of a state (the published mixing constants).
__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Backward sampling: in reverse elimination order each variable is drawn
// from P(var | later variables) proportional to exp(T_b), whose other scope
// variables are already drawn.  The read's seed is mixed into both words of
// its xorshift128+ state, so nearby seeds give unrelated streams.
__global__ __launch_bounds__(256) void tree_sample_kernel(TreeArgs a, const int* order_rev, int nbuckets, int n, int R,
                                                          const uint64_t* seeds, int8_t* samples, int8_t low_value) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    uint64_t s0 = splitmix(seeds[r] + 0x9e3779b97f4a7c15ull), s1 = splitmix(seeds[r] + 0x3c6ef372fe94f82aull);
    if ((s0 | s1) == 0) s0 = ~0ull;
    int8_t* x = samples + (size_t)r * n;
    for (int q = 0; q < nbuckets; ++q) {
        const TreeBucket B = a.buckets[order_rev[q]];
        int64_t idx = 0;
        for (int j = 1; j < B.k; ++j)
            if (x[a.scope[B.scope_off + j]] == 1) idx |= (int64_t)1 << j;
        const double t0 = a.T[B.t_off + idx], t1 = a.T[B.t_off + idx + 1];
        const double p1 = 1.0 / (1.0 + exp(t0 - t1));
        const double u = (double)(xs_next(s0, s1) >> 11) * 0x1.0p-53;
        x[B.var] = u < p1 ? (int8_t)1 : low_value;
    }
}

unsigned grid_for(int64_t work) {
    int64_t g = (work + 255) / 256;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

hipError_t launch_tree_forward(const TreeArgs& a, int b, int k, hipStream_t stream) {
    hipLaunchKernelGGL(tree_forward_kernel, dim3(grid_for((int64_t)1 << (k - 1))), dim3(256), 0, stream, a, b);
    return hipGetLastError();
}

hipError_t launch_tree_downward(const TreeArgs& a, int child, int r, hipStream_t stream) {
    hipLaunchKernelGGL(tree_downward_kernel, dim3(grid_for((int64_t)1 << r)), dim3(256), 0, stream, a, child);
    return hipGetLastError();
}

hipError_t launch_tree_var_marginals(const TreeArgs& a, int nbuckets, double* out_p1, hipStream_t stream) {
    if (nbuckets <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_var_marginal_kernel, dim3(nbuckets), dim3(256), 0, stream, a, out_p1);
    return hipGetLastError();
}

hipError_t launch_tree_pair_marginals(const TreeArgs& a, const int* pairs, int npairs, double* out,
                                      hipStream_t stream) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_pair_marginal_kernel, dim3(npairs), dim3(256), 0, stream, a, pairs, out);
    return hipGetLastError();
}

hipError_t launch_tree_sample(const TreeArgs& a, const int* order_rev, int nbuckets, int n, int R,
                              const uint64_t* seeds, int8_t* samples, int8_t low_value, hipStream_t stream) {
    if (R <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_sample_kernel, dim3((R + 255) / 256), dim3(256), 0, stream, a, order_rev, nbuckets, n, R,
                       seeds, samples, low_value);
    return hipGetLastError();
}

}  // namespace misa
