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

struct misa_tree_plan {
    int n = 0;
    double beta = 1.0, low = -1.0;
    double const_energy = 0.0;  // self-loops of spin problems (s * s = 1)
    int treewidth = 0;
    std::vector<int> order;     // elimination order (variables); bucket q eliminates order[q]
    std::vector<misa::TreeBucket> buckets;
    std::vector<misa::TreeTerm> terms;
    std::vector<int> scope, children, pos;
    std::vector<int> pairs_u, pairs_v, pair_desc;  // merged couplers; (bucket, pos u, pos v) triples
    int64_t t_entries = 0, half_entries = 0;
    // device
    misa::TreeBucket* d_buckets = nullptr;
    misa::TreeTerm* d_terms = nullptr;
    int *d_scope = nullptr, *d_children = nullptr, *d_pos = nullptr, *d_pairs = nullptr, *d_order_rev = nullptr;
    double *d_T = nullptr, *d_mu = nullptr, *d_delta = nullptr;
    bool forward_done = false;
    double log_z = 0.0;
    ~misa_tree_plan() {
        dev_free(d_buckets);
        dev_free(d_terms);
        dev_free(d_scope);
        dev_free(d_children);
        dev_free(d_pos);
        dev_free(d_pairs);
        dev_free(d_order_rev);
        dev_free(d_T);
        dev_free(d_mu);
        dev_free(d_delta);
    }
};

namespace {

// Greedy min-fill: repeatedly eliminate the variable whose neighbours need the
// fewest new edges to become a clique (ties: fewer neighbours, then lower
// index).
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
            if (best < 0 || fill[v] < fill[best] || (fill[v] == fill[best] && adj[v].size() < adj[best].size()))
                best = v;
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
        pl->pair_desc.push_back(q);
        pl->pair_desc.push_back(pu);
        pl->pair_desc.push_back(pv);
    }
    int64_t toff = 0, hoff = 0;
    for (int q = 0; q < n; ++q) {
        auto& B = pl->buckets[q];
        B.scope_off = (int)pl->scope.size();
        pl->scope.insert(pl->scope.end(), scopes[q].begin(), scopes[q].end());
        B.term_off = (int)pl->terms.size();
        B.nterms = (int)own[q].size();
        pl->terms.insert(pl->terms.end(), own[q].begin(), own[q].end());
        B.child_off = (int)pl->children.size();
        B.nchild = (int)kids[q].size();
        pl->children.insert(pl->children.end(), kids[q].begin(), kids[q].end());
        B.pos_off = (int)pl->pos.size();
        B.free_off = B.pos_off;
        if (B.parent >= 0) {
            const auto& ps = scopes[B.parent];
            std::vector<char> used(ps.size(), 0);
            for (int j = 1; j < B.k; ++j) {
                const int pj = (int)(std::find(ps.begin(), ps.end(), scopes[q][j]) - ps.begin());
                pl->pos.push_back(pj);
                used[pj] = 1;
            }
            B.free_off = (int)pl->pos.size();
            for (size_t pj = 0; pj < ps.size(); ++pj)
                if (!used[pj]) pl->pos.push_back((int)pj);
        }
        B.t_off = toff;
        toff += (int64_t)1 << B.k;
        B.mu_off = hoff;
        B.d_off = hoff;
        hoff += (int64_t)1 << (B.k - 1);
    }
    pl->t_entries = toff;
    pl->half_entries = hoff;
    std::vector<int> rev(pl->order.size());
    for (int q = 0; q < n; ++q) rev[q] = n - 1 - q;
    int rc;
    if ((rc = dev_upload(&pl->d_buckets, pl->buckets)) || (rc = dev_upload(&pl->d_terms, pl->terms)) ||
        (rc = dev_upload(&pl->d_scope, pl->scope)) || (rc = dev_upload(&pl->d_children, pl->children)) ||
        (rc = dev_upload(&pl->d_pos, pl->pos)) || (rc = dev_upload(&pl->d_pairs, pl->pair_desc)) ||
        (rc = dev_upload(&pl->d_order_rev, rev))) {
        std::string keep = g_err;
        delete pl;
        return fail(rc, keep);
    }
    if (hipMalloc(&pl->d_T, sizeof(double) * std::max<int64_t>(toff, 1)) != hipSuccess ||
        hipMalloc(&pl->d_mu, sizeof(double) * std::max<int64_t>(hoff, 1)) != hipSuccess ||
        hipMalloc(&pl->d_delta, sizeof(double) * std::max<int64_t>(hoff, 1)) != hipSuccess) {
        delete pl;
        return fail(MISA_ERR_DEVICE, "device memory for the bucket tables (" + std::to_string(toff) + " entries)");
    }
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

int misa_tree_pairs(const misa_tree_plan* pl, int32_t* u, int32_t* v) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    std::copy(pl->pairs_u.begin(), pl->pairs_u.end(), u);
    std::copy(pl->pairs_v.begin(), pl->pairs_v.end(), v);
    return MISA_OK;
}

namespace {

misa::TreeArgs tree_args(misa_tree_plan* pl) {
    misa::TreeArgs a{};
    a.buckets = pl->d_buckets;
    a.terms = pl->d_terms;
    a.scope = pl->d_scope;
    a.children = pl->d_children;
    a.pos = pl->d_pos;
    a.beta = pl->beta;
    a.low = pl->low;
    a.T = pl->d_T;
    a.mu = pl->d_mu;
    a.delta = pl->d_delta;
    return a;
}

// Upward pass and log Z (sum of the roots' messages, minus beta times the
// constant energy of spin self-loops).
int tree_forward(misa_tree_plan* pl, hipStream_t s) {
    if (pl->forward_done) return MISA_OK;
    const misa::TreeArgs a = tree_args(pl);
    for (int q = 0; q < pl->n; ++q) MISA_HIP(misa::launch_tree_forward(a, q, pl->buckets[q].k, s));
    double lz = -pl->beta * pl->const_energy;
    for (int q = 0; q < pl->n; ++q)
        if (pl->buckets[q].parent < 0) {
            double m = 0.0;
            MISA_HIP(hipMemcpyAsync(&m, pl->d_mu + pl->buckets[q].mu_off, sizeof(double), hipMemcpyDeviceToHost, s));
            MISA_HIP(hipStreamSynchronize(s));
            lz += m;
        }
    pl->log_z = lz;
    pl->forward_done = true;
    return MISA_OK;
}

}  // namespace

int misa_tree_marginals(misa_tree_plan* pl, double* log_z, double* var_p1, double* pair_p, void* stream) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc;
    if ((rc = tree_forward(pl, s))) return rc;
    if (log_z) *log_z = pl->log_z;
    if (!var_p1 && !pair_p) return MISA_OK;
    const misa::TreeArgs a = tree_args(pl);
    // downward pass: roots have delta = 0, children in reverse elimination order
    for (int q = pl->n - 1; q >= 0; --q) {
        const auto& B = pl->buckets[q];
        if (B.parent < 0) {
            MISA_HIP(hipMemsetAsync(pl->d_delta + B.d_off, 0, sizeof(double), s));  // +0.0
        } else {
            MISA_HIP(misa::launch_tree_downward(a, q, B.k - 1, s));
        }
    }
    double* d_out = nullptr;
    const int np = (int)pl->pairs_u.size();
    const size_t bytes = sizeof(double) * ((size_t)pl->n + 4 * (size_t)np + 1);
    MISA_HIP(hipMalloc(&d_out, bytes));
    hipError_t e = misa::launch_tree_var_marginals(a, pl->n, d_out, s);
    if (e == hipSuccess) e = misa::launch_tree_pair_marginals(a, pl->d_pairs, np, d_out + pl->n, s);
    std::vector<double> host((size_t)pl->n + 4 * (size_t)np);
    if (e == hipSuccess) e = hipMemcpyAsync(host.data(), d_out, sizeof(double) * host.size(), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_out);
    if (e != hipSuccess) return fail(MISA_ERR_DEVICE, std::string("tree marginals: ") + hipGetErrorString(e));
    if (var_p1) std::copy(host.begin(), host.begin() + pl->n, var_p1);
    if (pair_p) std::copy(host.begin() + pl->n, host.end(), pair_p);
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
