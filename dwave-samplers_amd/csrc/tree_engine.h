// Exact inference by bucket elimination (tree_api.cpp <-> tree_kernel.hip).
//
// Bucket b eliminates variable var: its scope is var plus its neighbours at
// elimination time, ordered by elimination position, var at bit 0 of every
// table index (bit j = 1: scope variable j takes its high value).  Tables hold
// log-potentials: T_b over the scope (the bucket's own terms and its
// children's messages), mu_b over the separator (scope without var, bit j-1 =
// scope variable j), delta_b the downward message over the same separator.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

struct TreeBucket {
    int var;
    int k;            // scope size
    int parent;       // bucket index, or -1
    int scope_off;    // scope variables: scope[scope_off .. + k)
    int term_off;     // own terms: terms[term_off .. + nterms)
    int nterms;
    int child_off;    // children: children[child_off .. + nchild) (bucket indices)
    int nchild;
    int pos_off;      // positions of this bucket's separator variables in the parent's scope (k - 1 ints)
    int free_off;     // parent scope positions not in this bucket's separator (parent k - (k - 1) ints)
    int64_t t_off;    // T: 2^k doubles
    int64_t mu_off;   // mu: 2^(k-1) doubles
    int64_t d_off;    // delta: 2^(k-1) doubles
};

// -beta * coef * x(pa) * x(pb); pb < 0: unary term -beta * coef * x(pa).
struct TreeTerm {
    int pa, pb;
    double coef;
};

struct TreeArgs {
    const TreeBucket* buckets;
    const TreeTerm* terms;
    const int* scope;
    const int* children;
    const int* pos;    // separator positions, see TreeBucket
    double beta;
    double low;        // value of a clear bit: -1 (spin) or 0 (binary)
    double* T;
    double* mu;
    double* delta;
};

hipError_t launch_tree_forward(const TreeArgs& a, int b, int k, hipStream_t stream);
hipError_t launch_tree_downward(const TreeArgs& a, int child, int r, hipStream_t stream);
hipError_t launch_tree_var_marginals(const TreeArgs& a, int nbuckets, double* out_p1, hipStream_t stream);
// pairs: (bucket, position of u, position of v) triples; out: 4 probabilities
// per pair, index bit(u) + 2 bit(v).
hipError_t launch_tree_pair_marginals(const TreeArgs& a, const int* pairs, int npairs, double* out,
                                      hipStream_t stream);
// Exact samples: buckets in reverse elimination order, one read per thread.
hipError_t launch_tree_sample(const TreeArgs& a, const int* order_rev, int nbuckets, int n, int R,
                              const uint64_t* seeds, int8_t* samples, int8_t low_value, hipStream_t stream);

}  // namespace misa
