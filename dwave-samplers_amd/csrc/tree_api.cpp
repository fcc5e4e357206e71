// C-ABI host side of exact inference by bucket elimination
// (include/misamplers.h).
//
// Counterpart of the reference's TreeDecompositionSampler core
// (dwave/samplers/tree/samplers.py:235-359 calling sample_bqm_wrapper,
// tree/sample.pyx, and its bucket-tree engine tree/src/include/sample.hpp:
// log partition function, variable and pair marginals, exact Boltzmann
// samples).  The host chooses the elimination order (min-fill when none is
// given, as samplers.py:319-321 does), eliminates symbolically to get every
// bucket's scope and the bucket tree, and the device evaluates the tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "host.h"
#include "misamplers.h"
#include "tree_engine.h"

using namespace misa_host;

// One launch of a pass: blocks over one bucket (bucket >= 0), or the items
// [off, off + nblocks) of the pass's list.
struct Level {
    int bucket;
    int off;
    int nblocks;
};

struct misa_tree_plan {
    int n = 0;
    double beta = 1.0, low = -1.0;
    double const_energy = 0.0;  // self-loops of spin problems (s * s = 1)
    int treewidth = 0;
    std::vector<int> order;     // elimination order (variables); bucket q eliminates order[q]
    std::vector<misa::TreeBucket> buckets;
    std::vector<misa::TreeTerm> terms;
    std::vector<int> scope, roots;
    std::vector<misa::TreeChild> children;
    std::vector<int> pairs_u, pairs_v;  // merged couplers
    // block work lists: upward pass by level, downward pass by depth, marginal chunks
    std::vector<misa::TreeItem> fwd_items, down_items, marg_items;
    std::vector<Level> fwd_levels, down_levels;
    std::vector<misa::TreeQuery> queries;
    int64_t t_entries = 0, half_entries = 0, partial_slots = 0;
    // device
    misa::TreeBucket* d_buckets = nullptr;
    misa::TreeTerm* d_terms = nullptr;
    misa::TreeItem *d_fwd_items = nullptr, *d_down_items = nullptr, *d_marg_items = nullptr;
    misa::TreeQuery* d_queries = nullptr;
    misa::TreeChild* d_children = nullptr;
    int *d_scope = nullptr, *d_roots = nullptr, *d_order_rev = nullptr;
    double *d_T = nullptr, *d_mu = nullptr, *d_delta = nullptr, *d_partial = nullptr;
    double* d_out = nullptr;  // variable marginals, pair marginals, log Z
    // the two passes as HIP graphs (kernel nodes only), captured on first use
    hipGraphExec_t g_forward = nullptr, g_marginals = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    misa_tree_stats stats{};
    bool forward_done = false;
    ~misa_tree_plan() {
        if (g_forward) (void)hipGraphExecDestroy(g_forward);
        if (g_marginals) (void)hipGraphExecDestroy(g_marginals);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        dev_free(d_buckets);
        dev_free(d_terms);
        dev_free(d_fwd_items);
        dev_free(d_down_items);
        dev_free(d_marg_items);
        dev_free(d_queries);
        dev_free(d_scope);
        dev_free(d_children);
        dev_free(d_roots);
        dev_free(d_order_rev);
        dev_free(d_T);
        dev_free(d_mu);
        dev_free(d_delta);
        dev_free(d_partial);
        dev_free(d_out);
    }
};

namespace {

// Greedy min-fill as the reference's heuristic scores it
// (tree/utilities.pyx:101-180): eliminate the variable with the fewest missing
// edges among its neighbours plus its neighbour count (that count is the
// pairs of a neighbour with itself, which the reference's scan includes);
// ties go to the lowest index (the reference takes the first in hash-map
// order).
std::vector<int> min_fill_order(int n, std::vector<std::set<int>> adj) {
    std::vector<int> order;
    order.reserve(n);
    std::vector<char> alive(n, 1);
    std::vector<long long> fill(n, 0);
    std::vector<char> dirty(n, 1);
    auto fill_of = [&](int v) {
        long long f = 0;
        for (auto i = adj[v].begin(); i != adj[v].end(); ++i) {
            auto j = i;
            for (++j; j != adj[v].end(); ++j)
                if (!adj[*i].count(*j)) ++f;
        }
        return f;
    };
    for (int step = 0; step < n; ++step) {
        int best = -1;
        for (int v = 0; v < n; ++v) {
            if (!alive[v]) continue;
            if (dirty[v]) {
                fill[v] = fill_of(v);
                dirty[v] = 0;
            }
            if (best < 0 || fill[v] + (long long)adj[v].size() < fill[best] + (long long)adj[best].size()) best = v;
        }
        order.push_back(best);
        alive[best] = 0;
        std::vector<int> nb(adj[best].begin(), adj[best].end());
        for (size_t i = 0; i < nb.size(); ++i)
            for (size_t j = i + 1; j < nb.size(); ++j) {
                adj[nb[i]].insert(nb[j]);
                adj[nb[j]].insert(nb[i]);
            }
        for (int u : nb) adj[u].erase(best);
        // fills change for the neighbours and their neighbours
        for (int u : nb) {
            dirty[u] = 1;
            for (int w : adj[u]) dirty[w] = 1;
        }
        adj[best].clear();
    }
    return order;
}

// One thread per separator index of each bucket of the level, 256 per block.
Level level_work(const misa_tree_plan* pl, const std::vector<int>& buckets, std::vector<misa::TreeItem>& items) {
    Level L{-1, (int)items.size(), 0};
    if (buckets.size() == 1) {
        L.bucket = buckets[0];
        L.nblocks = (int)((((int64_t)1 << (pl->buckets[L.bucket].k - 1)) + misa::TREE_THREADS - 1) /
                          misa::TREE_THREADS);
        return L;
    }
    for (int q : buckets)
        for (int64_t i = 0; i < ((int64_t)1 << (pl->buckets[q].k - 1)); i += misa::TREE_THREADS)
            items.push_back({q, (int)i});
    L.nblocks = (int)items.size() - L.off;
    return L;
}

}  // namespace

extern "C" {

int misa_tree_plan_create(const misa_problem* p, const int32_t* elimination_order, double beta, int low,
                          int max_treewidth, misa_tree_plan** out) {
    if (!p || !out) return fail(MISA_ERR_ARG, "null pointer");
    *out = nullptr;
    if (low != -1 && low != 0) return fail(MISA_ERR_ARG, "low must be -1 (spin) or 0 (binary)");
    if (!std::isfinite(beta)) return fail(MISA_ERR_ARG, "beta must be finite");
    max_treewidth = std::min(max_treewidth, 30);  // tables of 2^31 doubles at most
    const int n = p->n;
    auto pl = new misa_tree_plan();
    pl->n = n;
    pl->beta = beta;
    pl->low = low;
    // merged couplers in first-appearance order; self-loops are constants
    // (spin) or fields (binary)
    std::vector<double> h(p->h.begin(), p->h.end());
    std::map<std::pair<int, int>, int> index;
    std::vector<double> pw;
    for (int64_t c = 0; c < p->m; ++c) {
        int u = p->cs[c], v = p->ce[c];
        const double w = p->cw[c];
        if (u == v) {
            if (low == -1)
                pl->const_energy += w;
            else
                h[u] += w;
            continue;
        }
        auto key = std::make_pair(std::min(u, v), std::max(u, v));
        auto it = index.find(key);
        if (it == index.end()) {
            index.emplace(key, (int)pw.size());
            pl->pairs_u.push_back(u);
            pl->pairs_v.push_back(v);
            pw.push_back(w);
        } else {
            pw[it->second] += w;
        }
    }
    std::vector<std::set<int>> adj(n);
    for (size_t k = 0; k < pw.size(); ++k) {
        adj[pl->pairs_u[k]].insert(pl->pairs_v[k]);
        adj[pl->pairs_v[k]].insert(pl->pairs_u[k]);
    }
    if (elimination_order) {
        pl->order.assign(elimination_order, elimination_order + n);
        std::vector<char> seen(n, 0);
        for (int v : pl->order) {
            if (v < 0 || v >= n || seen[v]) {
                delete pl;
                return fail(MISA_ERR_ARG, "elimination order is not a permutation of the variables");
            }
            seen[v] = 1;
        }
    } else {
        pl->order = min_fill_order(n, adj);
    }
    std::vector<int> at(n);
    for (int q = 0; q < n; ++q) at[pl->order[q]] = q;
    // symbolic elimination: scopes ordered by elimination position
    std::vector<std::vector<int>> scopes(n);
    {
        std::vector<std::set<int>> g = adj;
        for (int q = 0; q < n; ++q) {
            const int v = pl->order[q];
            std::vector<int> nb(g[v].begin(), g[v].end());
            std::sort(nb.begin(), nb.end(), [&](int x, int y) { return at[x] < at[y]; });
            scopes[q].push_back(v);
            scopes[q].insert(scopes[q].end(), nb.begin(), nb.end());
            pl->treewidth = std::max(pl->treewidth, (int)nb.size());
            for (size_t i = 0; i < nb.size(); ++i)
                for (size_t j = i + 1; j < nb.size(); ++j) {
                    g[nb[i]].insert(nb[j]);
                    g[nb[j]].insert(nb[i]);
                }
            for (int u : nb) g[u].erase(v);
            g[v].clear();
        }
    }
    if (pl->treewidth > max_treewidth) {
        const int tw = pl->treewidth;
        delete pl;
        return fail(MISA_ERR_UNSUPPORTED, "maximum treewidth of " + std::to_string(max_treewidth) +
                                              " exceeded (elimination order width " + std::to_string(tw) + ")");
    }
    // buckets, children, positions, tables
    pl->buckets.resize(n);
    std::vector<std::vector<int>> kids(n);
    for (int q = 0; q < n; ++q) {
        auto& B = pl->buckets[q];
        B.var = pl->order[q];
        B.k = (int)scopes[q].size();
        B.parent = B.k > 1 ? at[scopes[q][1]] : -1;
        if (B.parent >= 0) kids[B.parent].push_back(q);
    }
    std::vector<std::vector<misa::TreeTerm>> own(n);
    for (int v = 0; v < n; ++v)
        if (h[v] != 0.0) own[at[v]].push_back({0, -1, h[v]});
    for (size_t k = 0; k < pw.size(); ++k) {
        const int u = pl->pairs_u[k], v = pl->pairs_v[k];
        const int q = std::min(at[u], at[v]);
        const auto& sc = scopes[q];
        const int pu = (int)(std::find(sc.begin(), sc.end(), u) - sc.begin());
        const int pv = (int)(std::find(sc.begin(), sc.end(), v) - sc.begin());
        own[q].push_back({pu, pv, pw[k]});
    }
    int64_t toff = 0, hoff = 0;
    for (int q = 0; q < n; ++q) {
        auto& B = pl->buckets[q];
        B.scope_off = (int)pl->scope.size();
        pl->scope.insert(pl->scope.end(), scopes[q].begin(), scopes[q].end());
        B.term_off = (int)pl->terms.size();
        B.nterms = (int)own[q].size();
        pl->terms.insert(pl->terms.end(), own[q].begin(), own[q].end());
        B.t_off = toff;
        toff += (int64_t)1 << B.k;
        B.mu_off = hoff;
        B.d_off = hoff;
        hoff += (int64_t)1 << (B.k - 1);
    }
    // separator positions in the parent's scope (increasing: both scopes are
    // ordered by elimination position) as masks, and the children's records
    for (int q = 0; q < n; ++q) {
        auto& B = pl->buckets[q];
        if (B.parent >= 0) {
            const auto& ps = scopes[B.parent];
            const auto& P = pl->buckets[B.parent];
            uint32_t used = 0;
            for (int j = 1; j < B.k; ++j)
                used |= 1u << (int)(std::find(ps.begin(), ps.end(), scopes[q][j]) - ps.begin());
            B.sep_mask = used;
            B.free_mask = (P.k >= 32 ? ~0u : ((1u << P.k) - 1)) & ~used;
            B.pk = P.k;
            B.pt_off = P.t_off;
            B.pd_off = P.d_off;
        }
    }
    for (int q = 0; q < n; ++q) {
        auto& B = pl->buckets[q];
        B.child_off = (int)pl->children.size();
        B.nchild = (int)kids[q].size();
        for (int c : kids[q]) {
            const auto& C = pl->buckets[c];
            pl->children.push_back({C.mu_off, C.sep_mask, C.k, c});
        }
    }
    pl->t_entries = toff;
    pl->half_entries = hoff;
    // roots (parents have higher indices than their children)
    for (int q = n - 1; q >= 0; --q) {
        auto& B = pl->buckets[q];
        B.root = B.parent < 0 ? q : pl->buckets[B.parent].root;
        if (B.parent < 0) pl->roots.push_back(q);
    }
    std::reverse(pl->roots.begin(), pl->roots.end());
    // upward levels: one more than the children's highest
    {
        std::vector<int> level(n, 0);
        int top = -1;
        for (int q = 0; q < n; ++q) {
            for (int c : kids[q]) level[q] = std::max(level[q], level[c] + 1);
            top = std::max(top, level[q]);
        }
        std::vector<std::vector<int>> by(top + 1);
        for (int q = 0; q < n; ++q) by[level[q]].push_back(q);
        for (const auto& l : by) pl->fwd_levels.push_back(level_work(pl, l, pl->fwd_items));
    }
    // downward depths below the roots
    {
        std::vector<int> depth(n, 0);
        int deepest = 0;
        for (int q = n - 1; q >= 0; --q)
            if (pl->buckets[q].parent >= 0) {
                depth[q] = depth[pl->buckets[q].parent] + 1;
                deepest = std::max(deepest, depth[q]);
            }
        std::vector<std::vector<int>> by(deepest + 1);
        for (int q = n - 1; q >= 0; --q) by[depth[q]].push_back(q);
        for (int d = 1; d <= deepest; ++d) pl->down_levels.push_back(level_work(pl, by[d], pl->down_items));
    }
    // marginal queries: each answered by the smallest bucket whose scope holds
    // its variables (every bucket's T + delta is the exact scope marginal)
    {
        std::vector<std::vector<int>> holds(n);  // variable -> buckets whose scope holds it
        std::vector<std::vector<int>> sorted_scope(n);
        for (int q = 0; q < n; ++q) {
            for (int v : scopes[q]) holds[v].push_back(q);
            sorted_scope[q] = scopes[q];
            std::sort(sorted_scope[q].begin(), sorted_scope[q].end());
        }
        auto position = [&](int q, int v) {
            const auto& sc = scopes[q];
            return (int)(std::find(sc.begin(), sc.end(), v) - sc.begin());
        };
        std::vector<std::vector<misa::TreeQuery>> per(n);
        for (int v = 0; v < n; ++v) {
            int best = at[v];
            for (int q : holds[v])
                if (pl->buckets[q].k < pl->buckets[best].k) best = q;
            per[best].push_back({best, position(best, v), -1, 0, 0, v});
        }
        for (size_t k = 0; k < pl->pairs_u.size(); ++k) {
            const int u = pl->pairs_u[k], v = pl->pairs_v[k];
            const auto& cand = holds[u].size() <= holds[v].size() ? holds[u] : holds[v];
            const int other = holds[u].size() <= holds[v].size() ? v : u;
            int best = std::min(at[u], at[v]);
            for (int q : cand)
                if (pl->buckets[q].k < pl->buckets[best].k &&
                    std::binary_search(sorted_scope[q].begin(), sorted_scope[q].end(), other))
                    best = q;
            per[best].push_back({best, position(best, u), position(best, v), 0, 0, n + 4 * (int)k});
        }
        for (int q = 0; q < n; ++q) {
            auto& B = pl->buckets[q];
            B.q_off = (int)pl->queries.size();
            B.nq = (int)per[q].size();
            if (per[q].empty()) continue;
            const int64_t size = (int64_t)1 << B.k;
            const int nchunks = (int)((size + misa::TREE_CHUNK - 1) / misa::TREE_CHUNK);
            for (auto Q : per[q]) {
                Q.slot = (int)pl->partial_slots;
                Q.nchunks = nchunks;
                pl->partial_slots += nchunks;
                pl->queries.push_back(Q);
            }
            for (int64_t e = 0; e < size; e += misa::TREE_CHUNK) pl->marg_items.push_back({q, (int)e});
        }
    }
    std::vector<int> rev(pl->order.size());
    for (int q = 0; q < n; ++q) rev[q] = n - 1 - q;
    int rc;
    if ((rc = dev_upload(&pl->d_buckets, pl->buckets)) || (rc = dev_upload(&pl->d_terms, pl->terms)) ||
        (rc = dev_upload(&pl->d_scope, pl->scope)) || (rc = dev_upload(&pl->d_children, pl->children)) ||
        (rc = dev_upload(&pl->d_roots, pl->roots)) ||
        (rc = dev_upload(&pl->d_fwd_items, pl->fwd_items)) || (rc = dev_upload(&pl->d_down_items, pl->down_items)) ||
        (rc = dev_upload(&pl->d_marg_items, pl->marg_items)) || (rc = dev_upload(&pl->d_queries, pl->queries)) ||
        (rc = dev_upload(&pl->d_order_rev, rev))) {
        std::string keep = g_err;
        delete pl;
        return fail(rc, keep);
    }
    const size_t nout = (size_t)n + 4 * pl->pairs_u.size() + 1;
    if (hipMalloc(&pl->d_T, sizeof(double) * std::max<int64_t>(toff, 1)) != hipSuccess ||
        hipMalloc(&pl->d_mu, sizeof(double) * std::max<int64_t>(hoff, 1)) != hipSuccess ||
        hipMalloc(&pl->d_delta, sizeof(double) * std::max<int64_t>(hoff, 1)) != hipSuccess ||
        hipMalloc(&pl->d_partial, sizeof(double) * 4 * std::max<int64_t>(pl->partial_slots, 1)) != hipSuccess ||
        hipMalloc(&pl->d_out, sizeof(double) * nout) != hipSuccess) {
        delete pl;
        return fail(MISA_ERR_DEVICE, "device memory for the bucket tables (" + std::to_string(toff) + " entries)");
    }
    // the roots' downward messages (empty separators) are 0; nothing writes them later
    if (hipMemset(pl->d_delta, 0, sizeof(double) * std::max<int64_t>(hoff, 1)) != hipSuccess ||
        hipMemset(pl->d_out, 0, sizeof(double) * nout) != hipSuccess) {
        delete pl;
        return fail(MISA_ERR_DEVICE, "clearing the downward messages");
    }
    for (auto& e : pl->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete pl;
            return fail(MISA_ERR_DEVICE, "tree plan events");
        }
    pl->stats.forward_levels = (int)pl->fwd_levels.size();
    pl->stats.downward_levels = (int)pl->down_levels.size();
    // algorithmic bytes: each table entry written or read once per use
    for (int q = 0; q < n; ++q) {
        const auto& B = pl->buckets[q];
        const double t = 8.0 * (double)((int64_t)1 << B.k), hlf = t / 2;
        pl->stats.forward_bytes += t + hlf;  // T and mu written
        for (int c = 0; c < B.nchild; ++c)
            pl->stats.forward_bytes += 4.0 * (double)((int64_t)1 << pl->children[B.child_off + c].k);  // child mu
        if (B.parent >= 0)  // parent T and delta read, own mu read, own delta written
            pl->stats.downward_bytes += 12.0 * (double)((int64_t)1 << B.pk) + 2 * hlf;
        if (B.nq > 0) pl->stats.marginal_bytes += t + hlf;  // T and delta read
    }
    pl->stats.marginal_blocks = (int64_t)pl->marg_items.size();
    pl->stats.table_entries = toff;
    *out = pl;
    return MISA_OK;
}

void misa_tree_plan_destroy(misa_tree_plan* plan) { delete plan; }

int misa_tree_plan_info(const misa_tree_plan* pl, int* treewidth, int64_t* table_entries, int* num_pairs,
                        int32_t* elimination_order) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    if (treewidth) *treewidth = pl->treewidth;
    if (table_entries) *table_entries = pl->t_entries;
    if (num_pairs) *num_pairs = (int)pl->pairs_u.size();
    if (elimination_order) std::copy(pl->order.begin(), pl->order.end(), elimination_order);
    return MISA_OK;
}

int misa_tree_recompute(misa_tree_plan* pl) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    pl->forward_done = false;
    return MISA_OK;
}

int misa_tree_pairs(const misa_tree_plan* pl, int32_t* u, int32_t* v) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    std::copy(pl->pairs_u.begin(), pl->pairs_u.end(), u);
    std::copy(pl->pairs_v.begin(), pl->pairs_v.end(), v);
    return MISA_OK;
}

}  // extern "C"

namespace {

misa::TreeArgs tree_args(misa_tree_plan* pl) {
    misa::TreeArgs a{};
    a.buckets = pl->d_buckets;
    a.terms = pl->d_terms;
    a.scope = pl->d_scope;
    a.children = pl->d_children;
    a.beta = pl->beta;
    a.low = pl->low;
    a.T = pl->d_T;
    a.mu = pl->d_mu;
    a.delta = pl->d_delta;
    return a;
}

double* log_z_slot(misa_tree_plan* pl) { return pl->d_out + pl->n + 4 * pl->pairs_u.size(); }

// Records one pass's launches on a private stream and instantiates them.
template <class F>
hipError_t capture(hipGraphExec_t* exec, F&& launches) {
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    hipGraph_t g = nullptr;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t le = launches(cs);
        e = hipStreamEndCapture(cs, &g);
        if (le != hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    return e;
}

// Upward pass: one launch per level, then log Z into the output buffer.
hipError_t capture_forward(misa_tree_plan* pl) {
    return capture(&pl->g_forward, [pl](hipStream_t cs) {
        const misa::TreeArgs a = tree_args(pl);
        for (const Level& L : pl->fwd_levels) {
            hipError_t e;
            const misa::TreeBucket* B = L.bucket >= 0 ? &pl->buckets[L.bucket] : nullptr;
            if (B && B->nterms <= misa::TREE_VALUE_TERMS && B->nchild <= misa::TREE_VALUE_CHILDREN) {
                // one bucket: its record travels as the kernel argument
                misa::TreeLevelArg V{};
                V.t_off = B->t_off;
                V.mu_off = B->mu_off;
                V.k = B->k;
                V.nterms = B->nterms;
                V.nchild = B->nchild;
                for (int t = 0; t < B->nterms; ++t) V.terms[t] = pl->terms[B->term_off + t];
                for (int c = 0; c < B->nchild; ++c) V.children[c] = pl->children[B->child_off + c];
                e = misa::launch_tree_forward_value(a, V, L.nblocks, cs);
            } else {
                e = misa::launch_tree_forward(a, L.bucket < 0 ? pl->d_fwd_items + L.off : nullptr, L.nblocks,
                                              L.bucket, cs);
            }
            if (e != hipSuccess) return e;
        }
        return misa::launch_tree_log_z(a, pl->d_roots, (int)pl->roots.size(), pl->const_energy, log_z_slot(pl), cs);
    });
}

// Downward pass: one launch per depth, then the marginal chunks.
hipError_t capture_marginals(misa_tree_plan* pl) {
    return capture(&pl->g_marginals, [pl](hipStream_t cs) {
        const misa::TreeArgs a = tree_args(pl);
        for (const Level& L : pl->down_levels) {
            hipError_t e;
            if (L.bucket >= 0) {
                // one bucket: its record travels as the kernel argument
                const misa::TreeBucket& C = pl->buckets[L.bucket];
                misa::TreeDownArg D{};
                D.pt_off = C.pt_off;
                D.pd_off = C.pd_off;
                D.d_off = C.d_off;
                D.mu_off = C.mu_off;
                D.sep_mask = C.sep_mask;
                D.free_mask = C.free_mask;
                D.k = C.k;
                e = misa::launch_tree_downward_value(a, D, L.nblocks, cs);
            } else {
                e = misa::launch_tree_downward(a, pl->d_down_items + L.off, L.nblocks, L.bucket, cs);
            }
            if (e != hipSuccess) return e;
        }
        return misa::launch_tree_marginals(a, pl->d_marg_items, (int)pl->marg_items.size(), pl->d_queries,
                                           (int)pl->queries.size(), pl->d_partial, pl->d_out, cs);
    });
}

int tree_forward(misa_tree_plan* pl, hipStream_t s) {
    pl->stats.forward_ms = 0.0;
    if (pl->forward_done) return MISA_OK;
    if (!pl->g_forward) MISA_HIP(capture_forward(pl));
    MISA_HIP(hipEventRecord(pl->ev[0], s));
    MISA_HIP(hipGraphLaunch(pl->g_forward, s));
    MISA_HIP(hipEventRecord(pl->ev[1], s));
    MISA_HIP(hipEventSynchronize(pl->ev[1]));
    float ms = 0.f;
    MISA_HIP(hipEventElapsedTime(&ms, pl->ev[0], pl->ev[1]));
    pl->stats.forward_ms = ms;
    pl->forward_done = true;
    return MISA_OK;
}

}  // namespace

extern "C" {

int misa_tree_marginals(misa_tree_plan* pl, double* log_z, double* var_p1, double* pair_p, void* stream) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc;
    if ((rc = tree_forward(pl, s))) return rc;
    pl->stats.marginals_ms = 0.0;
    const size_t np = pl->pairs_u.size();
    const bool marginals = (var_p1 || pair_p) && pl->n > 0;
    if (marginals) {
        if (!pl->g_marginals) MISA_HIP(capture_marginals(pl));
        MISA_HIP(hipEventRecord(pl->ev[1], s));
        MISA_HIP(hipGraphLaunch(pl->g_marginals, s));
        MISA_HIP(hipEventRecord(pl->ev[2], s));
    }
    std::vector<double> host((size_t)pl->n + 4 * np + 1);
    MISA_HIP(hipMemcpyAsync(host.data(), pl->d_out, sizeof(double) * host.size(), hipMemcpyDeviceToHost, s));
    MISA_HIP(hipStreamSynchronize(s));
    if (marginals) {
        float ms = 0.f;
        MISA_HIP(hipEventElapsedTime(&ms, pl->ev[1], pl->ev[2]));
        pl->stats.marginals_ms = ms;
    }
    if (log_z) *log_z = host[pl->n + 4 * np];
    if (var_p1) std::copy(host.begin(), host.begin() + pl->n, var_p1);
    if (pair_p) std::copy(host.begin() + pl->n, host.begin() + pl->n + 4 * np, pair_p);
    return MISA_OK;
}

int misa_tree_last_stats(const misa_tree_plan* pl, misa_tree_stats* out) {
    if (!pl || !out) return fail(MISA_ERR_ARG, "null pointer");
    *out = pl->stats;
    return MISA_OK;
}

int misa_tree_solve(misa_tree_plan* pl, int max_solutions, int8_t* samples, double* energies, int* num_found) {
    if (!pl || !num_found) return fail(MISA_ERR_ARG, "null pointer");
    *num_found = 0;
    if (max_solutions < 0) return fail(MISA_ERR_ARG, "max_solutions must be non-negative");
    const int n = pl->n;
    int64_t states = n >= 62 ? ((int64_t)1 << 62) : ((int64_t)1 << n);
    const int K = (int)std::min<int64_t>(max_solutions, states);
    if (K == 0 || n == 0) return MISA_OK;
    if ((!samples || !energies)) return fail(MISA_ERR_ARG, "null output buffer");
    if (K > (1 << 20)) return fail(MISA_ERR_UNSUPPORTED, "at most 2^20 solutions per call");
    int maxchild = 0;
    int64_t widest = 1;
    for (const auto& B : pl->buckets) {
        maxchild = std::max(maxchild, B.nchild);
        widest = std::max<int64_t>(widest, (int64_t)1 << (B.k - 1));
    }
    const int nb = (int)pl->buckets.size();
    const int nroots = (int)pl->roots.size();
    // per-bucket lists and back-pointers, per-thread scratch (forward in
    // chunks of at most 256 MiB of scratch), the components' back-pointers
    const int64_t rows = pl->half_entries;
    const int64_t fstride = misa::kbest_forward_scratch(K);
    const int64_t bstride = misa::kbest_backtrack_scratch(K, nb, maxchild);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(widest, ((int64_t)256 << 20) / fstride));
    const int64_t scr_bytes = std::max({chunk * fstride, (int64_t)K * bstride, 32 * (int64_t)K + 16});
    double* d_val = nullptr;
    uint32_t* d_bp = nullptr;
    int *d_len = nullptr, *d_rbi = nullptr, *d_rbj = nullptr, *d_total = nullptr;
    unsigned char* d_scr = nullptr;
    int8_t* d_samples = nullptr;
    double* d_en = nullptr;
    auto cleanup = [&]() {
        dev_free(d_val);
        dev_free(d_bp);
        dev_free(d_len);
        dev_free(d_rbi);
        dev_free(d_rbj);
        dev_free(d_total);
        dev_free(d_scr);
        dev_free(d_samples);
        dev_free(d_en);
    };
    if (hipMalloc(&d_val, sizeof(double) * rows * K) != hipSuccess ||
        hipMalloc(&d_bp, sizeof(uint32_t) * rows * K) != hipSuccess ||
        hipMalloc(&d_len, sizeof(int) * rows) != hipSuccess ||
        hipMalloc(&d_rbi, sizeof(int) * (int64_t)std::max(nroots, 1) * K) != hipSuccess ||
        hipMalloc(&d_rbj, sizeof(int) * (int64_t)std::max(nroots, 1) * K) != hipSuccess ||
        hipMalloc(&d_total, sizeof(int)) != hipSuccess || hipMalloc(&d_scr, scr_bytes) != hipSuccess ||
        hipMalloc(&d_samples, (size_t)K * n) != hipSuccess || hipMalloc(&d_en, sizeof(double) * K) != hipSuccess) {
        cleanup();
        return fail(MISA_ERR_DEVICE, "device memory for " + std::to_string(K) + " solutions over " +
                                         std::to_string(rows) + " table rows");
    }
    misa::KbArgs a{};
    a.buckets = pl->d_buckets;
    a.terms = pl->d_terms;
    a.children = pl->d_children;
    a.scope = pl->d_scope;
    a.low = pl->low;
    a.K = K;
    a.val = d_val;
    a.bp = d_bp;
    a.len = d_len;
    a.scr = d_scr;
    hipError_t e = hipSuccess;
    // buckets in elimination order: every child before its parent
    a.scr_stride = fstride;
    for (int q = 0; q < nb && e == hipSuccess; ++q) {
        const int64_t nsep = (int64_t)1 << (pl->buckets[q].k - 1);
        for (int64_t s0 = 0; s0 < nsep && e == hipSuccess; s0 += chunk)
            e = misa::launch_kbest_forward(a, q, s0, std::min(chunk, nsep - s0), nullptr);
    }
    if (e == hipSuccess)
        e = misa::launch_kbest_roots(a, pl->d_roots, nroots, pl->const_energy, d_rbi, d_rbj, d_en, d_total, nullptr);
    a.scr_stride = bstride;
    if (e == hipSuccess)
        e = misa::launch_kbest_backtrack(a, pl->d_roots, nroots, d_rbi, d_rbj, d_total, n, nb, maxchild, d_samples,
                                         (int8_t)pl->low, nullptr);
    int total = 0;
    if (e == hipSuccess) e = hipMemcpy(&total, d_total, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && total > 0) e = hipMemcpy(samples, d_samples, (size_t)total * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && total > 0) e = hipMemcpy(energies, d_en, sizeof(double) * total, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(MISA_ERR_DEVICE, std::string("k-best solve: ") + hipGetErrorString(e));
    *num_found = total;
    return MISA_OK;
}

int misa_tree_sample(misa_tree_plan* pl, int num_samples, const uint64_t* read_seeds, int8_t* samples, void* stream) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    if (num_samples == 0 || pl->n == 0) return MISA_OK;
    if (!read_seeds || !samples) return fail(MISA_ERR_ARG, "null buffer");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc;
    if ((rc = tree_forward(pl, s))) return rc;
    uint64_t* d_seeds = nullptr;
    int8_t* d_x = nullptr;
    const size_t total = (size_t)num_samples * pl->n;
    hipError_t e = hipMalloc(&d_seeds, sizeof(uint64_t) * num_samples);
    if (e == hipSuccess) e = hipMalloc(&d_x, total);
    if (e == hipSuccess) e = hipMemcpyAsync(d_seeds, read_seeds, sizeof(uint64_t) * num_samples, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = misa::launch_tree_sample(tree_args(pl), pl->d_order_rev, pl->n, pl->n, num_samples, d_seeds, d_x,
                                     (int8_t)pl->low, s);
    if (e == hipSuccess) e = hipMemcpyAsync(samples, d_x, total, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (d_seeds) (void)hipFree(d_seeds);
    if (d_x) (void)hipFree(d_x);
    if (e != hipSuccess) return fail(MISA_ERR_DEVICE, std::string("tree samples: ") + hipGetErrorString(e));
    return MISA_OK;
}

}  // extern "C"
