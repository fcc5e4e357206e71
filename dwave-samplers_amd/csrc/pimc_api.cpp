// C-ABI host side of the path-integral engine (include/misamplers.h).
//
// Replaces the reference's path-integral general_simulated_annealing
// (dwave/samplers/sqa/src/localPIMC.h:60-86, localPIMC.cpp:604-715): builds
// the adjacency without self-loops in coupler order, colours it, turns the
// (Hp, Hd) schedule into per-step slice tables, converts initial states /
// break lists to worldline words and back (localPIMC.cpp:169-199), launches
// the sweep and projection kernels and returns the completed read count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <string>
#include <vector>

#include "host.h"
#include "misamplers.h"
#include "pimc_engine.h"
#include "sa_engine.h"

using namespace misa_host;

struct misa_pimc_plan {
    misa_problem* p = nullptr;
    int n = 0, P = 32, spb = 1, nsteps = 0, kind = misa::PIMC_TABLE, nclass = 0, m2 = 0, ntab = 0;
    int maxdeg = 0, maxclass = 0;
    int Q = 0;  // chain updates: qubits per chain (0: site updates)
    int nchains = 0;
    unsigned maskP = 0xffffffffu;
    double J0 = 0.0, jchain = 0.0;
    std::vector<int> rowptr, class_off;
    std::vector<uint16_t> col, jidx, sites;
    std::vector<double> jtab, bs;
    std::vector<float> ginv;
    std::vector<int> cut;
    int *d_rowptr = nullptr, *d_coff = nullptr, *d_cut = nullptr;
    uint16_t *d_col = nullptr, *d_jidx = nullptr, *d_sites = nullptr;
    double *d_jtab = nullptr, *d_bs = nullptr;
    float* d_ginv = nullptr;
    ~misa_pimc_plan() {
        dev_free(d_rowptr); dev_free(d_coff); dev_free(d_cut); dev_free(d_col); dev_free(d_jidx);
        dev_free(d_sites); dev_free(d_jtab); dev_free(d_bs); dev_free(d_ginv);
    }
};

namespace {

// Per schedule step: Hp/P and the law of proposed cuts, P(cut) = tanh(Hd/P),
// as geometric gaps with ln 2 / ln(1 - tanh(Hd/P)).
void step_consts(double hp, double hd, int P, double& bs, float& ginv, int& cut) {
    bs = hp / (double)P;
    ginv = 0.0f;
    if (!(hd > 0.0)) {
        cut = misa::PIMC_CUT_NONE;
        return;
    }
    const double p = std::tanh(hd / (double)P);
    if (p >= 1.0) {
        cut = misa::PIMC_CUT_ALL;
        return;
    }
    const float g = (float)(0x1.62e42fefa39efp-1 / std::log1p(-p));
    if (!std::isfinite(g)) {  // p below ~1e-38: no cut is ever proposed
        cut = misa::PIMC_CUT_NONE;
        return;
    }
    cut = misa::PIMC_CUT_PROP;
    ginv = g;
}

// Worldlines resampled per sweep: every site, or Q passes over the chains of
// Q qubits.
int64_t qubit_updates_per_sweep(const misa_pimc_plan* pl) {
    return pl->Q > 0 ? (int64_t)pl->Q * pl->nchains * pl->Q : (int64_t)pl->n;
}

struct PimcGeo {
    int rpb_log2 = 0, threads = 64, lds = 0, bpc = 1, blocks = 0;
    int off_h = 0, off_jtab = 0, off_W = 0, off_rowptr = 0, off_coff = 0, off_col = 0, off_jidx = 0,
        off_sites = 0;
    int jtab_in_lds = 0;
};

void pimc_layout(const misa_pimc_plan* pl, int lg, PimcGeo& g) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 15) & ~(size_t)15;
        return (int)o;
    };
    const size_t n = (size_t)pl->n;
    g.rpb_log2 = lg;
    g.off_h = take(8 * n);
    const bool table = pl->kind == misa::PIMC_TABLE || pl->kind == misa::PIMC_CHAIN;
    g.jtab_in_lds = (table && (size_t)pl->ntab * 8 <= 16384) ? 1 : 0;
    g.off_jtab = take(g.jtab_in_lds ? 8 * (size_t)pl->ntab : 0);
    g.off_W = take((4 * n) << lg);
    g.off_rowptr = take(4 * (n + 1));
    g.off_coff = take(4 * ((size_t)pl->nclass + 1));
    g.off_col = take(2 * (size_t)pl->m2);
    g.off_jidx = take(table ? 2 * (size_t)pl->m2 : 0);
    g.off_sites = take(2 * n);
    g.lds = (int)off;
}

// Reads per block: the power of two minimising (rounds of blocks over the
// chip) x (per-thread task rounds summed over the colour classes).
int pimc_geometry(const misa_pimc_plan* pl, int R, PimcGeo& best) {
    const size_t lmax = misa::max_dynamic_lds();
    const int cus = device_cus();
    int forced = -1;
    if (const char* e = std::getenv("MISA_PIMC_READS_PER_BLOCK")) {
        const int v = std::atoi(e);
        for (int lg = 0; lg <= 6; ++lg)
            if ((1 << lg) == v) forced = lg;
    }
    double best_cost = -1.0;
    for (int lg = 0; lg <= 6; ++lg) {
        if (forced >= 0 && lg != forced) continue;
        PimcGeo g;
        pimc_layout(pl, lg, g);
        if ((size_t)g.lds > lmax) break;
        const int tasks = std::max(pl->maxclass, 1) << lg;
        g.threads = std::min(1024, ((tasks + 63) / 64) * 64);
        g.bpc = std::max(1, std::min((int)(lmax / (size_t)g.lds), 2048 / g.threads));
        g.blocks = (R + (1 << lg) - 1) >> lg;
        const double rounds = std::ceil((double)g.blocks / ((double)cus * g.bpc));
        double per = 0.0;
        for (int c = 0; c < pl->nclass; ++c) {
            const int t = (pl->class_off[c + 1] - pl->class_off[c]) << lg;
            per += std::ceil((double)t / g.threads);
        }
        const double cost = rounds * std::max(per, 1.0);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = g;
        }
    }
    if (best_cost < 0)
        return fail(MISA_ERR_UNSUPPORTED, "problem too large for the path-integral kernel's local memory (" +
                                              std::to_string(pl->n) + " variables, " + std::to_string(pl->m2) +
                                              " adjacency entries)");
    return MISA_OK;
}

// Break position of a kink between slices k-1 and k (k = 1..P, k = P being
// the wrap), midway between the slices' imaginary times k * 2^16 / P.
inline int kink_position(int k, int P) { return (int)(((int64_t)(2 * k - 1) * 65536) / (2 * P)); }

// Worldline word of a spin that starts (time 0) at s0 and flips at each of
// the given break positions (localPIMC.cpp:169-180 format): slice k takes the
// spin at time k * 2^16 / P.
uint32_t word_from_breaks(int s0, const int32_t* pos, int nb, int P) {
    uint32_t w = 0;
    for (int k = 0; k < P; ++k) {
        const int64_t t = ((int64_t)k * 65536) / P;
        int flips = 0;
        for (int j = 0; j < nb; ++j) flips += (int64_t)pos[j] <= t;
        const int s = (flips & 1) ? -s0 : s0;
        if (s > 0) w |= 1u << k;
    }
    return w;
}

// ---- Colour classes --------------------------------------------------------
// Sites of one class are pairwise non-adjacent, so a block updates a whole
// class at once; every class costs one barrier and ceil(tasks / threads)
// rounds of the block's threads, so few, equally sized classes are best.

// Distinct non-self neighbours per site.
void simple_adjacency(const std::vector<int>& rowptr, const std::vector<int>& adj, int n, std::vector<int>& off,
                      std::vector<int>& nb) {
    off.assign(n + 1, 0);
    nb.clear();
    std::vector<int> mark(n, -1);
    for (int v = 0; v < n; ++v) {
        for (int k = rowptr[v]; k < rowptr[v + 1]; ++k) {
            const int u = adj[k];
            if (u != v && mark[u] != v) {
                mark[u] = v;
                nb.push_back(u);
            }
        }
        off[v + 1] = (int)nb.size();
    }
}

int greedy_label_order(const std::vector<int>& off, const std::vector<int>& nb, int n, std::vector<int>& colour) {
    std::vector<int> mark(n + 1, -1);
    int nc = 0;
    for (int v = 0; v < n; ++v) {
        for (int k = off[v]; k < off[v + 1]; ++k)
            if (nb[k] < v) mark[colour[nb[k]]] = v;
        int c = 0;
        while (mark[c] == v) ++c;
        colour[v] = c;
        nc = std::max(nc, c + 1);
    }
    return nc;
}

// DSatur: colour next the uncoloured site with the most distinct neighbour
// colours (ties: most uncoloured neighbours, then lowest label).  kmax < 0:
// smallest admissible colour; otherwise the admissible colour below kmax with
// the fewest sites (ties: smallest), failing (returning -1) when none is left.
int dsatur(const std::vector<int>& off, const std::vector<int>& nb, int n, int kcap, int kmax,
           std::vector<int>& colour) {
    const int words = (kcap + 63) / 64;
    std::vector<uint64_t> used((size_t)n * words, 0);
    std::vector<int> sat(n, 0), udeg(n), size;
    std::fill(colour.begin(), colour.end(), -1);
    if (kmax >= 0) size.assign(kmax, 0);
    std::set<std::tuple<int, int, int>> queue;
    for (int v = 0; v < n; ++v) {
        udeg[v] = off[v + 1] - off[v];
        queue.emplace(0, -udeg[v], v);
    }
    int nc = 0;
    while (!queue.empty()) {
        const int v = std::get<2>(*queue.begin());
        queue.erase(queue.begin());
        const uint64_t* uv = &used[(size_t)v * words];
        int c = -1;
        if (kmax < 0) {
            for (c = 0; c < kcap && ((uv[c >> 6] >> (c & 63)) & 1); ++c) {
            }
            if (c >= kcap) return -1;
        } else {
            for (int t = 0; t < kmax; ++t)
                if (!((uv[t >> 6] >> (t & 63)) & 1) && (c < 0 || size[t] < size[c])) c = t;
            if (c < 0) return -1;
        }
        colour[v] = c;
        if ((int)size.size() <= c) size.resize(c + 1, 0);
        size[c]++;
        nc = std::max(nc, c + 1);
        for (int k = off[v]; k < off[v + 1]; ++k) {
            const int u = nb[k];
            if (colour[u] >= 0) continue;
            queue.erase(std::make_tuple(-sat[u], -udeg[u], u));
            uint64_t& w = used[(size_t)u * words + (c >> 6)];
            if (!((w >> (c & 63)) & 1)) {
                w |= 1ull << (c & 63);
                sat[u]++;
            }
            udeg[u]--;
            queue.emplace(-sat[u], -udeg[u], u);
        }
    }
    return nc;
}

// Rounds of 512 tasks summed over classes (about 2 reads of 1024 threads per
// block), then fewer classes.
double colouring_cost(const std::vector<int>& colour, int nc, int n) {
    std::vector<int> size(nc, 0);
    for (int v = 0; v < n; ++v) size[colour[v]]++;
    double rounds = 0.0;
    for (int c = 0; c < nc; ++c) rounds += std::ceil(size[c] / 512.0);
    return rounds * 1024.0 + nc;
}

// The engine's colouring: the cheapest of label-order greedy, DSatur, and
// balanced DSatur with DSatur's colour count (and one more).
int colour_sites(const std::vector<int>& rowptr, const std::vector<int>& adj, int n, std::vector<int>& colour) {
    std::vector<int> off, nb;
    simple_adjacency(rowptr, adj, n, off, nb);
    int maxdeg = 0;
    for (int v = 0; v < n; ++v) maxdeg = std::max(maxdeg, off[v + 1] - off[v]);
    std::vector<int> best(n, 0);
    int best_nc = greedy_label_order(off, nb, n, best);
    double best_cost = colouring_cost(best, best_nc, n);
    const int kcap = maxdeg + 2;
    if ((double)n * ((kcap + 63) / 64) * 8.0 > 64.0 * 1024 * 1024) {  // keep the bit sets modest
        colour = best;
        return best_nc;
    }
    std::vector<int> trial(n, 0);
    const int kstar = dsatur(off, nb, n, kcap, -1, trial);
    auto consider = [&](int nc) {
        if (nc <= 0) return;
        const double c = colouring_cost(trial, nc, n);
        if (c < best_cost) {
            best_cost = c;
            best_nc = nc;
            best = trial;
        }
    };
    consider(kstar);
    for (int k = kstar; k <= kstar + 1 && k <= kcap; ++k) consider(dsatur(off, nb, n, kcap, k, trial));
    colour = best;
    return best_nc;
}

// Chain colouring: chains c < n / Q are adjacent when a coupler joins a
// qubit of one to a qubit of the other (qubits in no chain are never updated
// and impose nothing); coloured as sites are.  Returns the class count.
int chain_colour(int n, int Q, const std::vector<int>& rowptr, const std::vector<int>& adj,
                 std::vector<int>& colour) {
    const int nch = n / Q;
    std::vector<int> crow(nch + 1, 0), cadj;
    for (int c = 0; c < nch; ++c) {
        for (int i = c * Q; i < c * Q + Q; ++i)
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
                const int d = adj[k] / Q;
                if (d != c && d < nch) cadj.push_back(d);
            }
        crow[c + 1] = (int)cadj.size();
    }
    if (cadj.empty()) cadj.push_back(0);
    colour.assign(std::max(nch, 1), 0);
    return nch > 0 ? colour_sites(crow, cadj, nch, colour) : 0;
}

// Adjacency without self-loops in coupler order (rowptr n + 1, adj).
void problem_adjacency(const misa_problem* p, std::vector<int>& rowptr, std::vector<int>& adj) {
    const int n = p->n;
    rowptr.assign(n + 1, 0);
    for (int64_t c = 0; c < p->m; ++c)
        if (p->cs[c] != p->ce[c]) {
            rowptr[p->cs[c] + 1]++;
            rowptr[p->ce[c] + 1]++;
        }
    for (int v = 0; v < n; ++v) rowptr[v + 1] += rowptr[v];
    adj.assign(std::max(rowptr[n], 1), 0);
    std::vector<int> fill(rowptr.begin(), rowptr.end());
    for (int64_t c = 0; c < p->m; ++c) {
        const int u = p->cs[c], v = p->ce[c];
        if (u == v) continue;
        adj[fill[u]++] = v;
        adj[fill[v]++] = u;
    }
}

// Chain colour classes of the plan: class_off / sites (chain ids, padded to n).
void colour_chains(misa_pimc_plan* pl, const std::vector<int>& adj) {
    const int n = pl->n, nch = pl->nchains;
    std::vector<int> colour;
    const int nc = chain_colour(n, pl->Q, pl->rowptr, adj, colour);
    pl->nclass = nc;
    pl->class_off.assign(nc + 1, 0);
    for (int c = 0; c < nch; ++c) pl->class_off[colour[c] + 1]++;
    for (int k = 0; k < nc; ++k) pl->class_off[k + 1] += pl->class_off[k];
    std::vector<int> fill(pl->class_off.begin(), pl->class_off.end());
    pl->sites.assign(std::max(n, 1), 0);
    for (int c = 0; c < nch; ++c) pl->sites[fill[colour[c]]++] = (uint16_t)c;
    for (int k = 0; k < nc; ++k) pl->maxclass = std::max(pl->maxclass, pl->class_off[k + 1] - pl->class_off[k]);
}

// Q == 0: site updates; Q >= 1: chain updates of Q qubits with coupling jchain.
int plan_create(const misa_problem* pc, int sweeps_per_beta, const double* hp_field, const double* hd_field,
                int num_fields, int num_slices, int Q, double jchain, misa_pimc_plan** out) {
    if (!out) return fail(MISA_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!pc) return fail(MISA_ERR_ARG, "null problem");
    if (sweeps_per_beta < 0) return fail(MISA_ERR_ARG, "sweeps_per_beta must be non-negative");
    if (num_fields < 0 || (num_fields > 0 && (!hp_field || !hd_field)))
        return fail(MISA_ERR_ARG, "bad field schedule");
    if (num_slices < 1 || num_slices > 32) return fail(MISA_ERR_ARG, "num_slices must be in [1, 32]");
    if (Q < 0) return fail(MISA_ERR_ARG, "qubits_per_chain must be positive");
    if (Q > misa::kChainQMax || Q * num_slices > misa::kChainNodes)
        return fail(MISA_ERR_UNSUPPORTED, "chain updates support up to " + std::to_string(misa::kChainQMax) +
                                              " qubits per chain and qubits_per_chain x num_slices <= " +
                                              std::to_string(misa::kChainNodes));
    misa_problem* p = const_cast<misa_problem*>(pc);
    const int n = p->n;
    auto* pl = new misa_pimc_plan();
    pl->p = p;
    pl->n = n;
    pl->P = num_slices;
    pl->maskP = num_slices == 32 ? 0xffffffffu : ((1u << num_slices) - 1u);
    pl->spb = sweeps_per_beta;
    pl->nsteps = num_fields;
    // Adjacency without self-loops, coupler order (localPIMC.cpp:652-674).
    pl->rowptr.assign(n + 1, 0);
    double J0 = -1.0;
    bool same = true;
    for (int64_t c = 0; c < p->m; ++c) {
        const int u = p->cs[c], v = p->ce[c];
        if (u == v) continue;
        pl->rowptr[u + 1]++;
        pl->rowptr[v + 1]++;
        const double a = std::fabs(p->cw[c]);
        if (J0 < 0) J0 = a;
        else if (a != J0) same = false;
    }
    for (int v = 0; v < n; ++v) pl->rowptr[v + 1] += pl->rowptr[v];
    const int64_t m2 = pl->rowptr[n];
    pl->kind = (same && J0 > 0 && n <= 32768) ? misa::PIMC_SIGNED : misa::PIMC_TABLE;
    if (Q > 0) {
        pl->kind = misa::PIMC_CHAIN;  // coupling table, float64 fields
        pl->Q = Q;
        pl->nchains = n / Q;
        pl->jchain = jchain;
    }
    if (n > 65535 || (pl->kind == misa::PIMC_SIGNED && n > 32767) || m2 > INT32_MAX / 4) {
        delete pl;
        return fail(MISA_ERR_UNSUPPORTED, "path-integral kernel supports up to 65535 variables");
    }
    pl->m2 = (int)m2;
    pl->J0 = J0 > 0 ? J0 : 0.0;
    pl->col.assign(std::max<int64_t>(m2, 1), 0);
    std::vector<double> Jrow(std::max<int64_t>(m2, 1), 0.0);
    std::vector<int> adj(std::max<int64_t>(m2, 1), 0);
    {
        std::vector<int> fill(pl->rowptr.begin(), pl->rowptr.end());
        for (int64_t c = 0; c < p->m; ++c) {
            const int u = p->cs[c], v = p->ce[c];
            if (u == v) continue;
            const double w = p->cw[c];
            int k = fill[u]++;
            adj[k] = v;
            Jrow[k] = w;
            k = fill[v]++;
            adj[k] = u;
            Jrow[k] = w;
        }
    }
    if (pl->kind == misa::PIMC_SIGNED) {
        for (int64_t k = 0; k < m2; ++k) pl->col[k] = (uint16_t)(adj[k] | (Jrow[k] < 0 ? 0x8000 : 0));
    } else {
        std::map<uint64_t, int> index;  // by bit pattern: keeps -0.0 / NaN payloads distinct
        pl->jidx.assign(std::max<int64_t>(m2, 1), 0);
        for (int64_t k = 0; k < m2; ++k) {
            pl->col[k] = (uint16_t)adj[k];
            uint64_t bits;
            std::memcpy(&bits, &Jrow[k], 8);
            auto it = index.find(bits);
            if (it == index.end()) {
                if (pl->jtab.size() >= 65535) {
                    delete pl;
                    return fail(MISA_ERR_UNSUPPORTED, "more than 65535 distinct coupling values");
                }
                it = index.emplace(bits, (int)pl->jtab.size()).first;
                pl->jtab.push_back(Jrow[k]);
            }
            pl->jidx[k] = (uint16_t)it->second;
        }
    }
    pl->ntab = (int)pl->jtab.size();
    for (int v = 0; v < n; ++v) pl->maxdeg = std::max(pl->maxdeg, pl->rowptr[v + 1] - pl->rowptr[v]);
    // Colour classes (colour order, ascending labels inside a class).
    if (Q > 0) {
        colour_chains(pl, adj);
    } else {
        std::vector<int> colour(n, 0);
        const int nc = colour_sites(pl->rowptr, adj, n, colour);
        pl->nclass = nc;
        pl->class_off.assign(nc + 1, 0);
        for (int v = 0; v < n; ++v) pl->class_off[colour[v] + 1]++;
        for (int c = 0; c < nc; ++c) pl->class_off[c + 1] += pl->class_off[c];
        std::vector<int> fill(pl->class_off.begin(), pl->class_off.end());
        pl->sites.assign(std::max(n, 1), 0);
        for (int v = 0; v < n; ++v) pl->sites[fill[colour[v]]++] = (uint16_t)v;
        for (int c = 0; c < nc; ++c) pl->maxclass = std::max(pl->maxclass, pl->class_off[c + 1] - pl->class_off[c]);
    }
    pl->bs.resize(std::max(num_fields, 1));
    pl->ginv.resize(std::max(num_fields, 1));
    pl->cut.resize(std::max(num_fields, 1));
    for (int k = 0; k < num_fields; ++k) step_consts(hp_field[k], hd_field[k], num_slices, pl->bs[k], pl->ginv[k], pl->cut[k]);
    int rc;
    if ((rc = p->ensure_device()) || (rc = dev_upload(&pl->d_rowptr, pl->rowptr)) ||
        (rc = dev_upload(&pl->d_coff, pl->class_off)) || (rc = dev_upload(&pl->d_col, pl->col)) ||
        (rc = dev_upload(&pl->d_sites, pl->sites)) || (rc = dev_upload(&pl->d_bs, pl->bs)) ||
        (rc = dev_upload(&pl->d_ginv, pl->ginv)) || (rc = dev_upload(&pl->d_cut, pl->cut))) {
        delete pl;
        return rc;
    }
    if ((pl->kind == misa::PIMC_TABLE || pl->kind == misa::PIMC_CHAIN) &&
        ((rc = dev_upload(&pl->d_jidx, pl->jidx)) || (rc = dev_upload(&pl->d_jtab, pl->jtab)))) {
        delete pl;
        return rc;
    }
    *out = pl;
    return MISA_OK;
}

}  // namespace

extern "C" {

int misa_pimc_plan_create(const misa_problem* p, int sweeps_per_beta, const double* hp_field,
                          const double* hd_field, int num_fields, int num_slices, misa_pimc_plan** out) {
    return plan_create(p, sweeps_per_beta, hp_field, hd_field, num_fields, num_slices, 0, 0.0, out);
}

int misa_pimc_plan_create_chains(const misa_problem* p, int sweeps_per_beta, const double* hp_field,
                                 const double* hd_field, int num_fields, int num_slices, int qubits_per_chain,
                                 double chain_coupler_strength, misa_pimc_plan** out) {
    if (qubits_per_chain < 1) {
        if (out) *out = nullptr;
        return fail(MISA_ERR_ARG, "qubits_per_chain must be positive");
    }
    return plan_create(p, sweeps_per_beta, hp_field, hd_field, num_fields, num_slices, qubits_per_chain,
                       chain_coupler_strength, out);
}

void misa_pimc_plan_destroy(misa_pimc_plan* plan) { delete plan; }

int misa_pimc_colouring(const misa_problem* p, int32_t* colour_out) {
    if (!p || (p->n > 0 && !colour_out)) return fail(MISA_ERR_ARG, "null argument");
    const int n = p->n;
    std::vector<int> rowptr, adj;
    problem_adjacency(p, rowptr, adj);
    std::vector<int> colour(n, 0);
    const int nc = colour_sites(rowptr, adj, n, colour);
    for (int v = 0; v < n; ++v) colour_out[v] = colour[v];
    return nc;
}

int misa_pimc_chain_colouring(const misa_problem* p, int qubits_per_chain, int32_t* colour_out) {
    if (!p) return fail(MISA_ERR_ARG, "null argument");
    if (qubits_per_chain < 1) return fail(MISA_ERR_ARG, "qubits_per_chain must be positive");
    const int nch = p->n / qubits_per_chain;
    if (nch > 0 && !colour_out) return fail(MISA_ERR_ARG, "null argument");
    std::vector<int> rowptr, adj, colour;
    problem_adjacency(p, rowptr, adj);
    const int nc = chain_colour(p->n, qubits_per_chain, rowptr, adj, colour);
    for (int c = 0; c < nch; ++c) colour_out[c] = colour[c];
    return nc;
}

int misa_pimc_plan_info(const misa_pimc_plan* pl, int num_reads, misa_pimc_stats* info) {
    if (!pl || !info) return fail(MISA_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    PimcGeo g;
    int rc = pimc_geometry(pl, std::max(num_reads, 1), g);
    if (rc) return rc;
    info->num_slices = pl->P;
    info->reads_per_block = 1 << g.rpb_log2;
    info->threads_per_block = g.threads;
    info->num_classes = pl->nclass;
    info->kind = pl->kind;
    info->lds_bytes = g.lds;
    info->max_degree = pl->maxdeg;
    info->blocks = g.blocks;
    info->site_updates = (int64_t)num_reads * pl->spb * pl->nsteps * qubit_updates_per_sweep(pl);
    return MISA_OK;
}

int misa_pimc_run_device(misa_pimc_plan* pl, uint32_t* d_W, const uint64_t* d_seeds, int R, int8_t* d_stats,
                         int interval, uint64_t* d_flips, void* stream) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    if (R < 0 || interval < 0) return fail(MISA_ERR_ARG, "negative size");
    if (R == 0 || pl->n == 0) return MISA_OK;
    if (!d_W || !d_seeds || (interval > 0 && !d_stats)) return fail(MISA_ERR_ARG, "null buffer");
    PimcGeo g;
    int rc = pimc_geometry(pl, R, g);
    if (rc) return rc;
    misa::PimcArgs a{};
    a.n = pl->n;
    a.R = R;
    a.P = pl->P;
    a.maskP = pl->maskP;
    a.nsteps = pl->nsteps;
    a.spb = pl->spb;
    a.interval = interval;
    a.ncol = interval > 0 ? 1 + (std::max(pl->nsteps, 1) - 1) / interval : 0;
    a.rpb_log2 = g.rpb_log2;
    a.nclass = pl->nclass;
    a.m2 = pl->m2;
    a.ntab = pl->ntab;
    a.jtab_in_lds = g.jtab_in_lds;
    a.J0 = pl->J0;
    a.Q = pl->Q;
    a.jchain = pl->jchain;
    a.h = pl->p->d_h;
    a.rowptr = pl->d_rowptr;
    a.col = pl->d_col;
    a.jidx = pl->d_jidx;
    a.jtab = pl->d_jtab;
    a.class_off = pl->d_coff;
    a.class_sites = pl->d_sites;
    a.bs = pl->d_bs;
    a.ginv = pl->d_ginv;
    a.cut = pl->d_cut;
    a.seeds = d_seeds;
    a.W = d_W;
    a.stats = interval > 0 ? d_stats : nullptr;
    a.flips = d_flips;
    a.off_h = g.off_h;
    a.off_jtab = g.off_jtab;
    a.off_W = g.off_W;
    a.off_rowptr = g.off_rowptr;
    a.off_coff = g.off_coff;
    a.off_col = g.off_col;
    a.off_jidx = g.off_jidx;
    a.off_sites = g.off_sites;
    a.lds_bytes = g.lds;
    a.threads = g.threads;
    // All |J| equal and rows of at most 15 neighbours: per-slice neighbour
    // counts as bit planes (same integer counts, so the same results).
    const int variant = (pl->kind == misa::PIMC_SIGNED && pl->maxdeg <= 15) ? misa::PIMC_PLANES : pl->kind;
    MISA_HIP(misa::launch_pimc(a, variant, reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

int misa_pimc_project_device(const misa_problem* pc, const uint32_t* d_W, int R, int8_t* d_states,
                             double* d_energies, void* stream) {
    if (!pc) return fail(MISA_ERR_ARG, "null problem");
    if (R < 0) return fail(MISA_ERR_ARG, "negative size");
    if (R == 0 || pc->n == 0) return MISA_OK;
    if (!d_W) return fail(MISA_ERR_ARG, "null buffer");
    misa_problem* p = const_cast<misa_problem*>(pc);
    int rc = p->ensure_device();
    if (rc) return rc;
    MISA_HIP(misa::launch_pimc_project(d_W, R, p->n, p->d_h, p->m, p->d_cs, p->d_ce, p->d_cw, d_states, d_energies,
                                       reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

int misa_pimc_sample(const misa_problem* p, int8_t* states, double* energies, int num_samples, int project_inputs,
                     int project_outputs, int32_t* num_breaks, const int32_t* breaks_in, int64_t breaks_in_size,
                     int32_t* breaks_buffer_out, int64_t breaks_buffer_size, int sweeps_per_beta,
                     const double* hp_field, const double* hd_field, int num_fields, double gamma,
                     double chain_coupler_strength, int qubits_per_chain, int qubits_per_update,
                     const uint64_t* read_seeds, int num_slices, int8_t* statistics, int interval,
                     misa_interrupt_cb interrupt, void* interrupt_ctx, misa_pimc_stats* stats) {
    (void)gamma;
    if (!p) return fail(MISA_ERR_ARG, "null problem");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    if (qubits_per_chain < 1) return fail(MISA_ERR_ARG, "qubits_per_chain must be positive");
    // localPIMC.cpp:101 asserts qubitsPerChain >= qubitsPerUpdate; any
    // qubits_per_update other than 1 selects chain updates (:115-126).
    if (qubits_per_chain < qubits_per_update)
        return fail(MISA_ERR_ARG, "qubits_per_chain must be at least qubits_per_update");
    const int chainQ = qubits_per_update != 1 ? qubits_per_chain : 0;
    if (interval < 0) return fail(MISA_ERR_ARG, "schedule_sample_interval must be non-negative");
    const int n = p->n;
    // pimc_annealing.pyx:197-203: nothing to sample.
    if ((int64_t)num_samples * n == 0) {
        for (int r = 0; r < num_samples; ++r) energies[r] = 0.0;
        return num_samples;
    }
    if (!states || !energies || !read_seeds) return fail(MISA_ERR_ARG, "null buffer");
    if (interval > 0 && !statistics) return fail(MISA_ERR_ARG, "null statistics buffer");
    if ((!project_inputs || !project_outputs) && !num_breaks) return fail(MISA_ERR_ARG, "null num_breaks");
    const size_t total = (size_t)num_samples * n;
    for (size_t i = 0; i < total; ++i)
        if (states[i] != 1 && states[i] != -1) return fail(MISA_ERR_ARG, "initial states must be +1/-1 spins");
    misa_pimc_plan* pl = nullptr;
    int rc = plan_create(p, sweeps_per_beta, hp_field, hd_field, num_fields, num_slices, chainQ,
                         chain_coupler_strength, &pl);
    if (rc) return rc;
    const int P = pl->P;
    // Initial worldlines (localPIMC.cpp:169-180).
    std::vector<uint32_t> W(total);
    if (project_inputs) {
        for (size_t i = 0; i < total; ++i) W[i] = states[i] > 0 ? pl->maskP : 0u;
    } else {
        int64_t cursor = 0;
        for (size_t i = 0; i < total; ++i) {
            const int nb = num_breaks[i];
            if (nb < 0 || cursor + nb > breaks_in_size || (nb > 0 && !breaks_in)) {
                delete pl;
                return fail(MISA_ERR_ARG, "breaks_in holds fewer positions than num_breaks adds up to");
            }
            W[i] = word_from_breaks(states[i], breaks_in + cursor, nb, P);
            cursor += nb;
        }
    }
    PimcGeo geo;
    if ((rc = pimc_geometry(pl, num_samples, geo))) {
        delete pl;
        return rc;
    }
    const int ncol = interval > 0 ? 1 + (std::max(num_fields, 1) - 1) / interval : 0;
    uint32_t* d_W = nullptr;
    uint64_t* d_seeds = nullptr;
    int8_t *d_stats = nullptr, *d_states = nullptr;
    double* d_en = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int done = 0;
    float total_ms = 0.f;
    auto cleanup = [&]() {
        if (d_W) (void)hipFree(d_W);
        if (d_seeds) (void)hipFree(d_seeds);
        if (d_stats) (void)hipFree(d_stats);
        if (d_states) (void)hipFree(d_states);
        if (d_en) (void)hipFree(d_en);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        delete pl;
    };
#define PIMC_TRY(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            cleanup();                                                                     \
            return fail(MISA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                  \
    } while (0)
    PIMC_TRY(hipMalloc(&d_W, sizeof(uint32_t) * total));
    PIMC_TRY(hipMalloc(&d_seeds, sizeof(uint64_t) * num_samples));
    PIMC_TRY(hipMalloc(&d_states, total));
    PIMC_TRY(hipMalloc(&d_en, sizeof(double) * num_samples));
    if (ncol) PIMC_TRY(hipMalloc(&d_stats, (size_t)num_samples * ncol * n));
    PIMC_TRY(hipMemcpy(d_W, W.data(), sizeof(uint32_t) * total, hipMemcpyHostToDevice));
    PIMC_TRY(hipMemcpy(d_seeds, read_seeds, sizeof(uint64_t) * num_samples, hipMemcpyHostToDevice));
    PIMC_TRY(hipEventCreate(&e0));
    PIMC_TRY(hipEventCreate(&e1));
    // Without an interrupt callback all reads go in one launch; with one, in
    // chunks of one full set of resident blocks, polling after every read.
    int chunk = num_samples;
    if (interrupt) chunk = std::max((geo.bpc * device_cus()) << geo.rpb_log2, 1);
    for (int r0 = 0; r0 < num_samples;) {
        const int R = std::min(chunk, num_samples - r0);
        PIMC_TRY(hipEventRecord(e0, nullptr));
        rc = misa_pimc_run_device(pl, d_W + (size_t)r0 * n, d_seeds + r0, R,
                                  ncol ? d_stats + (size_t)r0 * ncol * n : nullptr, interval, nullptr, nullptr);
        if (!rc)
            rc = misa_pimc_project_device(p, d_W + (size_t)r0 * n, R, d_states + (size_t)r0 * n, d_en + r0, nullptr);
        if (rc) {
            std::string keep = g_err;
            cleanup();
            return fail(rc, keep);
        }
        PIMC_TRY(hipEventRecord(e1, nullptr));
        PIMC_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        PIMC_TRY(hipEventElapsedTime(&ms, e0, e1));
        total_ms += ms;
        PIMC_TRY(hipMemcpy(states + (size_t)r0 * n, d_states + (size_t)r0 * n, (size_t)R * n, hipMemcpyDeviceToHost));
        PIMC_TRY(hipMemcpy(energies + r0, d_en + r0, sizeof(double) * R, hipMemcpyDeviceToHost));
        if (ncol)
            PIMC_TRY(hipMemcpy(statistics + (size_t)r0 * ncol * n, d_stats + (size_t)r0 * ncol * n,
                               (size_t)R * ncol * n, hipMemcpyDeviceToHost));
        if (!project_outputs)
            PIMC_TRY(hipMemcpy(W.data() + (size_t)r0 * n, d_W + (size_t)r0 * n, sizeof(uint32_t) * R * n,
                               hipMemcpyDeviceToHost));
        bool stop = false;
        for (int r = r0; r < r0 + R; ++r) {
            done = r + 1;
            if (interrupt && interrupt(interrupt_ctx)) {
                stop = true;
                break;
            }
        }
        r0 += R;
        if (stop) break;
    }
    // Final worldlines in the reference's break format and buffer discipline
    // (localPIMC.cpp:188-199, cursor advance :703-706).
    if (!project_outputs) {
        int64_t cursor = 0;
        for (int r = 0; r < done; ++r) {
            int64_t capacity = breaks_buffer_size - cursor;
            int64_t written = 0, read_total = 0;
            for (int i = 0; i < n; ++i) {
                const uint32_t w = W[(size_t)r * n + i];
                int nb = 0;
                int pos[32];
                for (int k = 1; k <= P; ++k) {
                    const uint32_t a = (w >> (k - 1)) & 1u, b = (w >> (k % P)) & 1u;
                    if (a != b) pos[nb++] = kink_position(k, P);
                }
                num_breaks[(size_t)r * n + i] = nb;
                if (capacity >= nb) {
                    for (int j = 0; j < nb; ++j) breaks_buffer_out[cursor + written + j] = pos[j];
                    written += nb;
                    capacity -= nb;
                }
                read_total += nb;
            }
            cursor += read_total;
        }
    }
    if (stats) {
        stats->kernel_ms = total_ms;
        stats->site_updates = (int64_t)done * pl->spb * pl->nsteps * qubit_updates_per_sweep(pl);
        stats->num_slices = P;
        stats->reads_per_block = 1 << geo.rpb_log2;
        stats->threads_per_block = geo.threads;
        stats->num_classes = pl->nclass;
        stats->kind = pl->kind;
        stats->lds_bytes = geo.lds;
        stats->max_degree = pl->maxdeg;
        stats->blocks = geo.blocks;
    }
    cleanup();
#undef PIMC_TRY
    return done;
}

}  // extern "C"
