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

// The bits of idx at the set positions of mask, packed from bit 0.
__device__ __forceinline__ int64_t extract_bits(int64_t idx, uint32_t mask) {
    int64_t j = 0;
    for (int t = 0; mask; ++t, mask &= mask - 1) j |= ((idx >> __builtin_ctz(mask)) & 1) << t;
    return j;
}

// Bits 0, 1, ... of j placed at the set positions of mask.
__device__ __forceinline__ int64_t deposit_bits(int64_t j, uint32_t mask) {
    int64_t idx = 0;
    for (int t = 0; mask; ++t, mask &= mask - 1) idx |= ((j >> t) & 1) << __builtin_ctz(mask);
    return idx;
}

__device__ __forceinline__ void block_work(const TreeItem* items, int bucket, int& b, int64_t& first) {
    if (items) {
        const TreeItem it = items[blockIdx.x];
        b = it.bucket;
        first = it.first;
    } else {
        b = bucket;
        first = (int64_t)blockIdx.x * TREE_THREADS;
    }
}

// T_b over the scope and mu_b = log-sum-exp over bit 0; one thread per
// separator index.
__global__ __launch_bounds__(TREE_THREADS) void tree_forward_kernel(TreeArgs a, const TreeItem* items, int bucket) {
    int b;
    int64_t first;
    block_work(items, bucket, b, first);
    const TreeBucket B = a.buckets[b];
    const int64_t i = first + threadIdx.x;
    if (i >= ((int64_t)1 << (B.k - 1))) return;
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
            const TreeChild ch = a.children[B.child_off + c];
            s += a.mu[ch.mu_off + extract_bits(idx, ch.mask)];
        }
        v[xv] = s;
        a.T[B.t_off + idx] = s;
    }
    const double m = v[0] > v[1] ? v[0] : v[1];
    a.mu[B.mu_off + i] = m + log(exp(v[0] - m) + exp(v[1] - m));
}

__global__ __launch_bounds__(TREE_THREADS) void tree_forward_value_kernel(TreeArgs a, TreeLevelArg L) {
    const int64_t i = (int64_t)blockIdx.x * TREE_THREADS + threadIdx.x;
    if (i >= ((int64_t)1 << (L.k - 1))) return;
    double v[2];
#pragma unroll
    for (int xv = 0; xv < 2; ++xv) {
        const int64_t idx = (i << 1) | xv;
        double s = 0.0;
        for (int t = 0; t < L.nterms; ++t) {
            const TreeTerm tm = L.terms[t];
            const double xa = ((idx >> tm.pa) & 1) ? 1.0 : a.low;
            const double xb = tm.pb < 0 ? 1.0 : (((idx >> tm.pb) & 1) ? 1.0 : a.low);
            s += -a.beta * tm.coef * xa * xb;
        }
        for (int c = 0; c < L.nchild; ++c) {
            const TreeChild ch = L.children[c];
            s += a.mu[ch.mu_off + extract_bits(idx, ch.mask)];
        }
        v[xv] = s;
        a.T[L.t_off + idx] = s;
    }
    const double m = v[0] > v[1] ? v[0] : v[1];
    a.mu[L.mu_off + i] = m + log(exp(v[0] - m) + exp(v[1] - m));
}

// delta_c over child c's separator: the parent's belief T_p + delta_p summed
// (log-sum-exp) over the parent positions outside the separator, minus mu_c.
__global__ __launch_bounds__(TREE_THREADS) void tree_downward_kernel(TreeArgs a, const TreeItem* items, int bucket) {
    int b;
    int64_t first;
    block_work(items, bucket, b, first);
    const TreeBucket C = a.buckets[b];
    const int64_t j = first + threadIdx.x;
    if (j >= ((int64_t)1 << (C.k - 1))) return;
    const int64_t base = deposit_bits(j, C.sep_mask);
    const int64_t nf = (int64_t)1 << __builtin_popcount(C.free_mask);
    Lse acc = lse_empty();
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t idx = base | deposit_bits(f, C.free_mask);
        acc.add(a.T[C.pt_off + idx] + a.delta[C.pd_off + (idx >> 1)]);
    }
    a.delta[C.d_off + j] = acc.value() - a.mu[C.mu_off + j];
}

__global__ __launch_bounds__(TREE_THREADS) void tree_downward_value_kernel(TreeArgs a, TreeDownArg D) {
    const int64_t j = (int64_t)blockIdx.x * TREE_THREADS + threadIdx.x;
    if (j >= ((int64_t)1 << (D.k - 1))) return;
    const int64_t base = deposit_bits(j, D.sep_mask);
    const int64_t nf = (int64_t)1 << __builtin_popcount(D.free_mask);
    Lse acc = lse_empty();
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t idx = base | deposit_bits(f, D.free_mask);
        acc.add(a.T[D.pt_off + idx] + a.delta[D.pd_off + (idx >> 1)]);
    }
    a.delta[D.d_off + j] = acc.value() - a.mu[D.mu_off + j];
}

// log Z: the roots' mu summed in a fixed order (strided per thread, then a
// fixed tree), minus beta times the constant energy.
__global__ __launch_bounds__(TREE_THREADS) void tree_log_z_kernel(TreeArgs a, const int* roots, int nroots,
                                                                  double const_energy, double* log_z) {
    __shared__ double sh[TREE_THREADS];
    double s = 0.0;
    for (int t = threadIdx.x; t < nroots; t += TREE_THREADS) s += a.mu[a.buckets[roots[t]].mu_off];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = TREE_THREADS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *log_z = sh[0] - a.beta * const_energy;
}

// Per chunk of a bucket's table: p = exp(T + delta - log Z of the component)
// once per entry, then for each query of the bucket the sums of p over its
// cells, reduced across the block in a fixed order.
__global__ __launch_bounds__(TREE_THREADS) void tree_marginal_chunk_kernel(TreeArgs a, const TreeItem* items,
                                                                           const TreeQuery* queries,
                                                                           double* partial) {
    __shared__ double red[TREE_THREADS / 64][4];
    const TreeItem it = items[blockIdx.x];
    const TreeBucket B = a.buckets[it.bucket];
    const double norm = a.mu[a.buckets[B.root].mu_off];
    const int64_t size = (int64_t)1 << B.k;
    double p[TREE_PER_THREAD];
#pragma unroll
    for (int i = 0; i < TREE_PER_THREAD; ++i) {
        const int64_t e = (int64_t)it.first + threadIdx.x + TREE_THREADS * i;
        p[i] = e < size ? exp(a.T[B.t_off + e] + a.delta[B.d_off + (e >> 1)] - norm) : 0.0;
    }
    const int chunk = it.first / TREE_CHUNK;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int q = B.q_off; q < B.q_off + B.nq; ++q) {
        const TreeQuery Q = queries[q];
        double c[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < TREE_PER_THREAD; ++i) {
            const int64_t e = (int64_t)it.first + threadIdx.x + TREE_THREADS * i;
            const int cell = (int)((e >> Q.pa) & 1) | (Q.pb >= 0 ? (int)((e >> Q.pb) & 1) << 1 : 0);
#pragma unroll
            for (int x = 0; x < 4; ++x) c[x] += cell == x ? p[i] : 0.0;
        }
#pragma unroll
        for (int x = 0; x < 4; ++x)
            for (int off = 32; off > 0; off >>= 1) c[x] += __shfl_xor(c[x], off);
        if (lane == 0)
            for (int x = 0; x < 4; ++x) red[wave][x] = c[x];
        __syncthreads();
        if (threadIdx.x < 4) {
            double s = 0.0;
            for (int w = 0; w < TREE_THREADS / 64; ++w) s += red[w][threadIdx.x];
            partial[4 * ((int64_t)Q.slot + chunk) + threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// One thread per query: its chunks' partial sums in chunk order, normalised.
__global__ __launch_bounds__(TREE_THREADS) void tree_marginal_finish_kernel(const TreeQuery* queries, int nqueries,
                                                                            const double* partial, double* out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nqueries) return;
    const TreeQuery Q = queries[q];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < Q.nchunks; ++c)
        for (int x = 0; x < 4; ++x) s[x] += partial[4 * ((int64_t)Q.slot + c) + x];
    if (Q.pb < 0) {
        out[Q.dest] = s[1] / (s[0] + s[1]);
    } else {
        const double tot = s[0] + s[1] + s[2] + s[3];
        for (int x = 0; x < 4; ++x) out[Q.dest + x] = s[x] / tot;
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
__global__ __launch_bounds__(TREE_THREADS) void tree_sample_kernel(TreeArgs a, const int* order_rev, int nbuckets, int n, int R,
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

// ---- k lowest energies (min-sum with k-best lists) -------------------------

struct HNode {
    double v;
    int i, j;
};

__device__ __forceinline__ bool hless(const HNode& x, const HNode& y) {
    return x.v < y.v || (x.v == y.v && (x.i < y.i || (x.i == y.i && x.j < y.j)));
}

__device__ void heap_push(HNode* h, int& hn, HNode x) {
    int c = hn++;
    while (c > 0) {
        const int p = (c - 1) >> 1;
        if (!hless(x, h[p])) break;
        h[c] = h[p];
        c = p;
    }
    h[c] = x;
}

__device__ HNode heap_pop(HNode* h, int& hn) {
    const HNode top = h[0];
    const HNode x = h[--hn];
    int c = 0;
    for (;;) {
        int m = 2 * c + 1;
        if (m >= hn) break;
        if (m + 1 < hn && hless(h[m + 1], h[m])) ++m;
        if (!hless(h[m], x)) break;
        h[c] = h[m];
        c = m;
    }
    if (hn > 0) h[c] = x;
    return top;
}

// C = the min(K, la * lb) smallest A[i] + B[j] (A, B ascending, la, lb >= 1),
// in lexicographic (value, i, j) order: best-first over pairs where (i, j)
// follows (i, j-1), and (i, 0) follows (i-1, 0), so every pair is reached
// once, after a pair that precedes it.  ci / cj (if set) receive i and j.
__device__ int kbest_sum(const double* A, int la, const double* B, int lb, int K, double* C, int* ci, int* cj,
                         HNode* h) {
    const int64_t pairs = (int64_t)la * lb;
    const int lc = (int)(pairs < K ? pairs : K);
    int hn = 0;
    heap_push(h, hn, HNode{A[0] + B[0], 0, 0});
    for (int t = 0; t < lc; ++t) {
        const HNode top = heap_pop(h, hn);
        C[t] = top.v;
        if (ci) {
            ci[t] = top.i;
            cj[t] = top.j;
        }
        if (top.j + 1 < lb) heap_push(h, hn, HNode{A[top.i] + B[top.j + 1], top.i, top.j + 1});
        if (top.j == 0 && top.i + 1 < la) heap_push(h, hn, HNode{A[top.i + 1] + B[0], top.i + 1, 0});
    }
    return lc;
}

// Energy of a bucket's own terms at scope assignment idx.
__device__ double own_energy(const KbArgs& a, const TreeBucket& B, int64_t idx) {
    double e = 0.0;
    for (int t = 0; t < B.nterms; ++t) {
        const TreeTerm tm = a.terms[B.term_off + t];
        const double xa = ((idx >> tm.pa) & 1) ? 1.0 : a.low;
        const double xb = tm.pb < 0 ? 1.0 : (((idx >> tm.pb) & 1) ? 1.0 : a.low);
        e += tm.coef * xa * xb;
    }
    return e;
}

// L for scope assignment idx: the k-best sum of the children's lists, child
// by child, plus the own energy; returns its length, in *cur.  With bi / bj
// set, child c's back-pointers go to bi[c * K ..], bj[c * K ..].
__device__ int kbest_scope(const KbArgs& a, const TreeBucket& B, int64_t idx, double*& cur, double*& nxt, HNode* h,
                           int* bi, int* bj) {
    const int K = a.K;
    int lc = 1;
    cur[0] = 0.0;
    for (int c = 0; c < B.nchild; ++c) {
        const TreeChild ch = a.children[B.child_off + c];
        const int64_t row = ch.mu_off + extract_bits(idx, ch.mask);
        lc = kbest_sum(cur, lc, a.val + row * K, a.len[row], K, nxt, bi ? bi + (int64_t)c * K : nullptr,
                       bj ? bj + (int64_t)c * K : nullptr, h);
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    const double e = own_energy(a, B, idx);
    for (int r = 0; r < lc; ++r) cur[r] = e + cur[r];
    return lc;
}

__global__ __launch_bounds__(TREE_THREADS) void tree_kbest_forward_kernel(KbArgs a, int b, int64_t s0, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t s = s0 + t;
    const TreeBucket B = a.buckets[b];
    const int K = a.K;
    unsigned char* base = a.scr + t * a.scr_stride;
    double* P = reinterpret_cast<double*>(base);
    double* Q = P + K;
    double* L0 = Q + K;
    HNode* h = reinterpret_cast<HNode*>(L0 + K);
    double* cur = P;
    double* nxt = Q;
    const int len0 = kbest_scope(a, B, s << 1, cur, nxt, h, nullptr, nullptr);
    for (int r = 0; r < len0; ++r) L0[r] = cur[r];
    cur = P;
    nxt = Q;
    const int len1 = kbest_scope(a, B, (s << 1) | 1, cur, nxt, h, nullptr, nullptr);
    // merge the two outcomes of the variable (ties: 0 first)
    const int64_t row = B.mu_off + s;
    double* ov = a.val + row * K;
    uint32_t* ob = a.bp + row * K;
    const int lout = len0 + len1 < K ? len0 + len1 : K;
    int i0 = 0, i1 = 0;
    for (int r = 0; r < lout; ++r) {
        if (i1 >= len1 || (i0 < len0 && !(cur[i1] < L0[i0]))) {
            ov[r] = L0[i0];
            ob[r] = (uint32_t)i0++;
        } else {
            ov[r] = cur[i1];
            ob[r] = 0x80000000u | (uint32_t)i1++;
        }
    }
    a.len[row] = lout;
}

__global__ void tree_kbest_roots_kernel(KbArgs a, const int* roots, int nroots, double const_energy, int* root_bi,
                                        int* root_bj, double* energies, int* total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int K = a.K;
    double* cur = reinterpret_cast<double*>(a.scr);
    double* nxt = cur + K;
    HNode* h = reinterpret_cast<HNode*>(nxt + K);
    int lc = 1;
    cur[0] = 0.0;
    for (int r = 0; r < nroots; ++r) {
        const int64_t row = a.buckets[roots[r]].mu_off;  // a root's separator is empty
        lc = kbest_sum(cur, lc, a.val + row * K, a.len[row], K, nxt, root_bi + (int64_t)r * K,
                       root_bj + (int64_t)r * K, h);
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
    for (int t = 0; t < lc; ++t) energies[t] = cur[t] + const_energy;
    *total = lc;
}

__global__ __launch_bounds__(TREE_THREADS) void tree_kbest_backtrack_kernel(KbArgs a, const int* roots, int nroots,
                                                                            const int* root_bi, const int* root_bj,
                                                                            const int* total, int n, int nbuckets,
                                                                            int maxchild, int8_t* samples,
                                                                            int8_t low_value) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *total) return;
    const int K = a.K;
    unsigned char* base = a.scr + (int64_t)t * a.scr_stride;
    double* P = reinterpret_cast<double*>(base);
    double* Q = P + K;
    HNode* h = reinterpret_cast<HNode*>(Q + K);
    int* rank = reinterpret_cast<int*>(h + K + 1);
    int* sbi = rank + nbuckets;
    int* sbj = sbi + (int64_t)maxchild * K;
    int jj = t;
    for (int r = nroots - 1; r >= 0; --r) {
        rank[roots[r]] = root_bj[(int64_t)r * K + jj];
        jj = root_bi[(int64_t)r * K + jj];
    }
    int8_t* x = samples + (size_t)t * n;
    for (int q = nbuckets - 1; q >= 0; --q) {  // parents before their children
        const TreeBucket B = a.buckets[q];
        int64_t s = 0;
        for (int j = 1; j < B.k; ++j)
            if (x[a.scope[B.scope_off + j]] == 1) s |= (int64_t)1 << (j - 1);
        const uint32_t w = a.bp[(B.mu_off + s) * K + rank[q]];
        const int xv = (int)(w >> 31);
        x[B.var] = xv ? (int8_t)1 : low_value;
        if (B.nchild > 0) {
            double* cur = P;
            double* nxt = Q;
            (void)kbest_scope(a, B, (s << 1) | xv, cur, nxt, h, sbi, sbj);
            int pos = (int)(w & 0x7fffffffu);
            for (int c = B.nchild - 1; c >= 0; --c) {
                rank[a.children[B.child_off + c].bucket] = sbj[(int64_t)c * K + pos];
                pos = sbi[(int64_t)c * K + pos];
            }
        }
    }
}

}  // namespace

hipError_t launch_kbest_forward(const KbArgs& a, int bucket, int64_t s0, int64_t count, hipStream_t stream) {
    if (count <= 0) return hipSuccess;
    const int64_t blocks = (count + TREE_THREADS - 1) / TREE_THREADS;
    hipLaunchKernelGGL(tree_kbest_forward_kernel, dim3((unsigned)blocks), dim3(TREE_THREADS), 0, stream, a, bucket,
                       s0, count);
    return hipGetLastError();
}

hipError_t launch_kbest_roots(const KbArgs& a, const int* roots, int nroots, double const_energy, int* root_bi,
                              int* root_bj, double* energies, int* total, hipStream_t stream) {
    hipLaunchKernelGGL(tree_kbest_roots_kernel, dim3(1), dim3(64), 0, stream, a, roots, nroots, const_energy, root_bi,
                       root_bj, energies, total);
    return hipGetLastError();
}

hipError_t launch_kbest_backtrack(const KbArgs& a, const int* roots, int nroots, const int* root_bi, const int* root_bj,
                                  const int* total, int n, int nbuckets, int maxchild, int8_t* samples,
                                  int8_t low_value, hipStream_t stream) {
    hipLaunchKernelGGL(tree_kbest_backtrack_kernel, dim3((a.K + TREE_THREADS - 1) / TREE_THREADS), dim3(TREE_THREADS),
                       0, stream, a, roots, nroots, root_bi, root_bj, total, n, nbuckets, maxchild, samples,
                       low_value);
    return hipGetLastError();
}

hipError_t launch_tree_forward(const TreeArgs& a, const TreeItem* items, int nblocks, int bucket, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_forward_kernel, dim3(nblocks), dim3(TREE_THREADS), 0, stream, a, items, bucket);
    return hipGetLastError();
}

hipError_t launch_tree_forward_value(const TreeArgs& a, const TreeLevelArg& L, int nblocks, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_forward_value_kernel, dim3(nblocks), dim3(TREE_THREADS), 0, stream, a, L);
    return hipGetLastError();
}

hipError_t launch_tree_downward(const TreeArgs& a, const TreeItem* items, int nblocks, int bucket,
                                hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_downward_kernel, dim3(nblocks), dim3(TREE_THREADS), 0, stream, a, items, bucket);
    return hipGetLastError();
}

hipError_t launch_tree_downward_value(const TreeArgs& a, const TreeDownArg& D, int nblocks, hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_downward_value_kernel, dim3(nblocks), dim3(TREE_THREADS), 0, stream, a, D);
    return hipGetLastError();
}

hipError_t launch_tree_log_z(const TreeArgs& a, const int* roots, int nroots, double const_energy, double* log_z,
                             hipStream_t stream) {
    hipLaunchKernelGGL(tree_log_z_kernel, dim3(1), dim3(TREE_THREADS), 0, stream, a, roots, nroots, const_energy,
                       log_z);
    return hipGetLastError();
}

hipError_t launch_tree_marginals(const TreeArgs& a, const TreeItem* items, int nitems, const TreeQuery* queries,
                                 int nqueries, double* partial, double* out, hipStream_t stream) {
    if (nitems <= 0 || nqueries <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_marginal_chunk_kernel, dim3(nitems), dim3(TREE_THREADS), 0, stream, a, items, queries,
                       partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(tree_marginal_finish_kernel, dim3((nqueries + TREE_THREADS - 1) / TREE_THREADS),
                       dim3(TREE_THREADS), 0, stream, queries, nqueries, partial, out);
    return hipGetLastError();
}

hipError_t launch_tree_sample(const TreeArgs& a, const int* order_rev, int nbuckets, int n, int R,
                              const uint64_t* seeds, int8_t* samples, int8_t low_value, hipStream_t stream) {
    if (R <= 0) return hipSuccess;
    hipLaunchKernelGGL(tree_sample_kernel, dim3((R + TREE_THREADS - 1) / TREE_THREADS), dim3(TREE_THREADS), 0, stream, a, order_rev, nbuckets, n, R,
                       seeds, samples, low_value);
    return hipGetLastError();
}

}  // namespace misa
