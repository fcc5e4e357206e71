// Exact inference by bucket elimination (tree_api.cpp <-> tree_kernel.hip).
//
// Bucket b eliminates variable var: its scope is var plus its neighbours at
// elimination time, ordered by elimination position, var at bit 0 of every
// table index (bit j = 1: scope variable j takes its high value).  Tables hold
// log-potentials: T_b over the scope (the bucket's own terms and its
// children's messages), mu_b over the separator (scope without var, bit j-1 =
// scope variable j), delta_b the downward message over the same separator.
// T_b + delta_b is the log of the unnormalised joint marginal of the scope;
// it sums (log-sum-exp) to mu of the bucket's root, the log partition
// function of its connected component.
//
// Work is cut into blocks of 256 threads described by TreeItem lists: the
// upward pass runs one launch per level (a bucket's level is one more than
// its children's highest), the downward pass one launch per depth below the
// roots, and the marginals one launch over chunks of TREE_CHUNK entries of
// the buckets that answer a query.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

constexpr int TREE_THREADS = 256;
constexpr int TREE_PER_THREAD = 16;
constexpr int TREE_CHUNK = TREE_THREADS * TREE_PER_THREAD;   // table entries per marginal block

struct TreeBucket {
    int var;
    int k;            // scope size
    int parent;       // bucket index, or -1
    int root;         // bucket index of the root of this bucket's tree
    int scope_off;    // scope variables: scope[scope_off .. + k)
    int term_off;     // own terms: terms[term_off .. + nterms)
    int nterms;
    int child_off;    // children: children[child_off .. + nchild)
    int nchild;
    int q_off;        // marginal queries answered from this bucket: queries[q_off .. + nq)
    int nq;
    uint32_t sep_mask;   // parent scope positions of this bucket's separator variables
    uint32_t free_mask;  // parent scope positions outside the separator
    int pk;              // the parent's scope size and tables (copied here: one load fewer)
    int64_t pt_off, pd_off;
    int64_t t_off;    // T: 2^k doubles
    int64_t mu_off;   // mu: 2^(k-1) doubles
    int64_t d_off;    // delta: 2^(k-1) doubles
};

// A child's message as its parent reads it: mu[mu_off + the parent index's
// bits at the set positions of mask, packed from bit 0].  Scopes are ordered
// by elimination position, so the positions increase with the child's
// separator index bits.
struct TreeChild {
    int64_t mu_off;
    uint32_t mask;
    int k;
    int bucket;  // the child's bucket index
};

// -beta * coef * x(pa) * x(pb); pb < 0: unary term -beta * coef * x(pa).
struct TreeTerm {
    int pa, pb;
    double coef;
};

// One block's work: a bucket and the first table (or separator) index of its
// range.
struct TreeItem {
    int bucket;
    int first;
};

// A marginal: scope positions pa (and pb; pb < 0 for a single variable) of
// the bucket that answers it.  Its per-chunk partial sums (4 doubles, cell
// bit(pa) + 2 bit(pb)) sit at partial[4 (slot + chunk)]; the result goes to
// out[dest] (variable: P(high)) or out[dest .. + 4) (pair: the 4 cells).
struct TreeQuery {
    int bucket;
    int pa, pb;
    int slot;
    int nchunks;
    int dest;
};

struct TreeArgs {
    const TreeBucket* buckets;
    const TreeTerm* terms;
    const int* scope;
    const TreeChild* children;
    double beta;
    double low;        // value of a clear bit: -1 (spin) or 0 (binary)
    double* T;
    double* mu;
    double* delta;
};

// A one-bucket level's record passed by value (kernel argument), so the
// launch reads no bucket, term or child record from device memory first.
constexpr int TREE_VALUE_TERMS = 16;
constexpr int TREE_VALUE_CHILDREN = 8;
struct TreeLevelArg {
    int64_t t_off, mu_off;
    int k, nterms, nchild, pad;
    TreeTerm terms[TREE_VALUE_TERMS];
    TreeChild children[TREE_VALUE_CHILDREN];
};
// Upward messages of a one-bucket level from its record (nterms and nchild
// within the bounds above), the same operations in the same order as
// launch_tree_forward.
hipError_t launch_tree_forward_value(const TreeArgs& a, const TreeLevelArg& L, int nblocks, hipStream_t stream);

// A one-bucket depth's record for the downward pass, passed by value.
struct TreeDownArg {
    int64_t pt_off, pd_off, d_off, mu_off;
    uint32_t sep_mask, free_mask;
    int k, pad;
};
// Downward messages of a one-bucket depth from its record, the same
// operations in the same order as launch_tree_downward.
hipError_t launch_tree_downward_value(const TreeArgs& a, const TreeDownArg& D, int nblocks, hipStream_t stream);

// Upward messages of one level: blocks items[0 .. nblocks) (first =
// separator index), or, with items NULL, nblocks blocks over the one bucket
// of the level.
hipError_t launch_tree_forward(const TreeArgs& a, const TreeItem* items, int nblocks, int bucket, hipStream_t stream);
// Downward messages of one depth, the same way (first = separator index of
// the child).
hipError_t launch_tree_downward(const TreeArgs& a, const TreeItem* items, int nblocks, int bucket,
                                hipStream_t stream);
// log Z = sum of the roots' mu (fixed order) - beta * const_energy, to *log_z.
hipError_t launch_tree_log_z(const TreeArgs& a, const int* roots, int nroots, double const_energy, double* log_z,
                             hipStream_t stream);
// Partial sums of every query over the chunks items[0 .. nitems) (first =
// first table entry of the chunk), then the normalised results.
hipError_t launch_tree_marginals(const TreeArgs& a, const TreeItem* items, int nitems, const TreeQuery* queries,
                                 int nqueries, double* partial, double* out, hipStream_t stream);
// Exact samples: buckets in reverse elimination order, one read per thread.
hipError_t launch_tree_sample(const TreeArgs& a, const int* order_rev, int nbuckets, int n, int R,
                              const uint64_t* seeds, int8_t* samples, int8_t low_value, hipStream_t stream);

// k lowest energies (min-sum elimination with k-best lists).  For bucket b
// and separator assignment s, val[(mu_off + s) * K + r], r < len[mu_off + s],
// are the len smallest energies of b's eliminated subtree (its own terms and
// its descendants') given s, ascending; bp holds (x << 31) | j: the value of
// b's variable and the position in the list L_x of the scope assignment
// (s, x).  L_x is the k-best sum of the children's lists (own energy added),
// built child by child; a k-best sum keeps the smallest (value, i, j) in
// lexicographic order, so every list and tie order is reproducible.
struct KbArgs {
    const TreeBucket* buckets;
    const TreeTerm* terms;
    const TreeChild* children;
    const int* scope;
    double low;
    int K;
    double* val;
    uint32_t* bp;
    int* len;
    unsigned char* scr;   // per-thread scratch
    int64_t scr_stride;   // bytes per thread
};

// Per-thread scratch bytes: three lists and a heap (forward), or two lists,
// a heap, the bucket ranks and the children's back-pointers (backtrack).
inline int64_t kbest_forward_scratch(int K) { return ((24 * (int64_t)K + 16 * ((int64_t)K + 1)) + 15) / 16 * 16; }
inline int64_t kbest_backtrack_scratch(int K, int nbuckets, int maxchild) {
    return ((16 * (int64_t)K + 16 * ((int64_t)K + 1) + 4 * (int64_t)nbuckets + 8 * (int64_t)maxchild * K) + 15) / 16 *
           16;
}
// One bucket's lists for separator assignments [s0, s0 + count).
hipError_t launch_kbest_forward(const KbArgs& a, int bucket, int64_t s0, int64_t count, hipStream_t stream);
// The components' lists combined (one thread): energies[0 .. *total) ascending,
// plus const_energy; root_bi / root_bj [nroots * K] the back-pointers.
hipError_t launch_kbest_roots(const KbArgs& a, const int* roots, int nroots, double const_energy, int* root_bi,
                              int* root_bj, double* energies, int* total, hipStream_t stream);
// State t < *total, one thread each (K threads launched): buckets from the
// last eliminated down.
hipError_t launch_kbest_backtrack(const KbArgs& a, const int* roots, int nroots, const int* root_bi, const int* root_bj,
                                  const int* total, int n, int nbuckets, int maxchild, int8_t* samples,
                                  int8_t low_value, hipStream_t stream);

}  // namespace misa
