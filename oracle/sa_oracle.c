/*
 * ORACLE — test infrastructure only.
 *
 * Plain-C restatement of the reference simulated-annealing core
 * (dwave/samplers/sa/src/cpu_sa.cpp), used as the checker for the HIP engine
 * and as the "port" CPU baseline.  Nothing in the product links or calls this
 * file; only tests/, the smoke check and bench.py's cpu_baseline leg load it.
 *
 * Parity pins (see tests/test_oracle.py):
 *   - orc_sa_one / orc_sa_sample / orc_sa_sample_stream: bit-exact against
 *     golden vectors written by the compiled reference (oracle/_ref) and, where
 *     that library is present, against it directly; plus closed-form special
 *     cases (null BQM oscillation, beta=inf greedy limit, 0 sweeps identity).
 *   - orc_xorshift_next: known-answer values of xorshift128+ computed from the
 *     published recurrence in an independent Python big-int implementation.
 *   - orc_read_seeds: splitmix64 known-answer values (published test vector
 *     for seed 0: 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, ...).
 *
 * Per-read seeding: the reference seeds one xorshift128+ stream per call and
 * lets it run on across reads (cpu_sa.cpp:283-284, loop :321-358).  The GPU
 * engine runs reads in parallel, so parity is defined read by read: read r is
 * the reference called with num_samples=1, seed = read_seed[r] and its own
 * initial state.  orc_sa_sample_stream keeps the call-level semantics.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_GIBBS 0      /* cpu_sa.h:36-39 enum Proposal { Gibbs, Metropolis } */
#define ORC_METROPOLIS 1
#define ORC_SEQUENTIAL 0 /* cpu_sa.h:40-43 enum VariableOrder { Sequential, Random } */
#define ORC_RANDOM 1

/* xorshift128+ step, cpu_sa.cpp:27-34 (state pair s[0], s[1]). */
static inline uint64_t xs_next(uint64_t s[2]) {
    uint64_t a = s[0];
    const uint64_t b = s[1];
    s[0] = b;
    a ^= a << 23;
    s[1] = a ^ b ^ (a >> 17) ^ (b >> 26);
    return s[1] + b;
}

uint64_t orc_xorshift_next(uint64_t* s) { return xs_next(s); }

/* The r-th (0-based) output of splitmix64 started from `seed`.  This is the
 * engine's per-read seed derivation (not a reference function). */
void orc_read_seeds(uint64_t seed, int64_t R, uint64_t* out) {
    for (int64_t r = 0; r < R; ++r) {
        uint64_t z = seed + (uint64_t)(r + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        out[r] = z ^ (z >> 31);
    }
}

/* Adjacency in coupler insertion order, cpu_sa.cpp:298-316: for coupler c,
 * append v to u's list, then u to v's list (a self-loop appends twice).
 * Returns 0, or -1 for mismatched lengths is handled by the caller, -2 for an
 * out-of-range index (cpu_sa.cpp:302-304). */
int orc_build_csr(int n, int64_t m, const int* cs, const int* ce,
                  const double* cw, int64_t* rowptr, int* col, double* J) {
    for (int i = 0; i <= n; ++i) rowptr[i] = 0;
    for (int64_t c = 0; c < m; ++c) {
        int u = cs[c], v = ce[c];
        if (u < 0 || v < 0 || u >= n || v >= n) return -2;
        rowptr[u + 1]++;
        rowptr[v + 1]++;
    }
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) fill[i] = rowptr[i];
    for (int64_t c = 0; c < m; ++c) {
        int u = cs[c], v = ce[c];
        col[fill[u]] = v; J[fill[u]] = cw[c]; fill[u]++;
        col[fill[v]] = u; J[fill[v]] = cw[c]; fill[v]++;
    }
    free(fill);
    return 0;
}

/* get_state_energy, cpu_sa.cpp:210-227: linear terms in variable order, then
 * couplers in input order, each product formed as (s_u * J) * s_v. */
double orc_state_energy(int n, const int8_t* s, const double* h, int64_t m,
                        const int* cs, const int* ce, const double* cw) {
    double e = 0.0;
    for (int v = 0; v < n; ++v) e += s[v] * h[v];
    for (int64_t c = 0; c < m; ++c) e += s[cs[c]] * cw[c] * s[ce[c]];
    return e;
}

typedef struct {
    int n;
    const double* h;
    const int64_t* rowptr;
    const int* col;
    const double* J;
    int sweeps_per_beta;
    const double* betas;
    int nbetas;
    int order;
    int proposal;
} orc_problem;

/* simulated_annealing_run<order, proposal>, cpu_sa.cpp:93-197, with the RNG
 * state passed in (the reference keeps it thread_local, :41).  delta[] is
 * scratch of n doubles.  Optionally counts flips and the summed degree of
 * flipped variables (the engine's roofline bookkeeping). */
static void orc_run(const orc_problem* P, int8_t* s, double* delta, uint64_t rng[2],
                    uint64_t* flips, uint64_t* degsum) {
    const int n = P->n;
    /* get_flip_energy, cpu_sa.cpp:54-74: field summed in adjacency order. */
    for (int v = 0; v < n; ++v) {
        double e = P->h[v];
        for (int64_t k = P->rowptr[v]; k < P->rowptr[v + 1]; ++k)
            e += s[P->col[k]] * P->J[k];
        delta[v] = -2 * s[v] * e;
    }
    uint64_t nflip = 0, ndeg = 0;
    for (int b = 0; b < P->nbetas; ++b) {
        const double beta = P->betas[b];
        for (int sw = 0; sw < P->sweeps_per_beta; ++sw) {
            const double threshold = 44.36142 / beta;           /* :131 */
            for (int it = 0; it < n; ++it) {
                int v;
                if (P->order == ORC_RANDOM) {                   /* :134-136 */
                    uint64_t r = xs_next(rng);
                    v = (int)(r % (uint64_t)n);
                } else {
                    v = it;
                }
                if (delta[v] >= threshold) continue;            /* :140 */
                int flip = 0;
                if (P->proposal == ORC_METROPOLIS) {            /* :144-157 */
                    if (delta[v] <= 0.0) {
                        flip = 1;
                    } else {
                        uint64_t r = xs_next(rng);
                        if (exp(-delta[v] * beta) * (double)UINT64_MAX > (double)r) flip = 1;
                    }
                } else {                                        /* :158-166 */
                    uint64_t r = xs_next(rng);
                    if ((double)UINT64_MAX > (double)r * (1 + exp(delta[v] * beta))) flip = 1;
                }
                if (flip) {                                     /* :168-191 */
                    const int8_t mult = (int8_t)(4 * s[v]);
                    for (int64_t k = P->rowptr[v]; k < P->rowptr[v + 1]; ++k) {
                        int u = P->col[k];
                        delta[u] += mult * P->J[k] * s[u];
                    }
                    s[v] = (int8_t)(-s[v]);
                    delta[v] = -delta[v];
                    nflip++;
                    ndeg += (uint64_t)(P->rowptr[v + 1] - P->rowptr[v]);
                }
            }
        }
    }
    if (flips) *flips = nflip;
    if (degsum) *degsum = ndeg;
}

/* One read: reference semantics of general_simulated_annealing called with
 * num_samples=1 (cpu_sa.cpp:253-362): rng = {seed ? seed : UINT64_MAX, 0}. */
double orc_sa_one(int n, const double* h, const int64_t* rowptr, const int* col,
                  const double* J, int64_t m, const int* cs, const int* ce,
                  const double* cw, int sweeps_per_beta, const double* betas,
                  int nbetas, uint64_t seed, int order, int proposal, int8_t* state,
                  uint64_t* flips, uint64_t* degsum) {
    orc_problem P = {n, h, rowptr, col, J, sweeps_per_beta, betas, nbetas, order, proposal};
    uint64_t rng[2] = {seed ? seed : UINT64_MAX, 0};
    double* delta = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    orc_run(&P, state, delta, rng, flips, degsum);
    free(delta);
    return orc_state_energy(n, state, h, m, cs, ce, cw);
}

typedef struct {
    const orc_problem* P;
    int64_t m;
    const int* cs; const int* ce; const double* cw;
    const uint64_t* seeds;
    int8_t* states;
    double* energies;
    uint64_t* flips;
    uint64_t* degsum;
    int64_t r0, r1;
} orc_job;

static void* orc_worker(void* arg) {
    orc_job* jb = (orc_job*)arg;
    const orc_problem* P = jb->P;
    double* delta = (double*)malloc(sizeof(double) * (size_t)(P->n > 0 ? P->n : 1));
    for (int64_t r = jb->r0; r < jb->r1; ++r) {
        uint64_t rng[2] = {jb->seeds[r] ? jb->seeds[r] : UINT64_MAX, 0};
        int8_t* s = jb->states + r * P->n;
        uint64_t f = 0, d = 0;
        orc_run(P, s, delta, rng, &f, &d);
        if (jb->flips) jb->flips[r] = f;
        if (jb->degsum) jb->degsum[r] = d;
        jb->energies[r] = orc_state_energy(P->n, s, P->h, jb->m, jb->cs, jb->ce, jb->cw);
    }
    free(delta);
    return NULL;
}

/* Read-by-read sampling over `nthreads` POSIX threads (reads are independent
 * chains, so the result does not depend on nthreads).  Returns 0, -1 for
 * mismatched lengths, -2 for a bad index (cpu_sa.cpp:276-279, :302-304). */
int orc_sa_sample(int n, const double* h, int64_t m_s, int64_t m_e, int64_t m_w,
                  const int* cs, const int* ce, const double* cw, int sweeps_per_beta,
                  const double* betas, int nbetas, const uint64_t* seeds, int64_t R,
                  int8_t* states, double* energies, int order, int proposal,
                  int nthreads, uint64_t* flips, uint64_t* degsum) {
    if (!(m_s == m_e && m_s == m_w)) return -1;
    const int64_t m = m_s;
    int64_t* rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int* col = (int*)malloc(sizeof(int) * (size_t)(2 * m + 1));
    double* J = (double*)malloc(sizeof(double) * (size_t)(2 * m + 1));
    int rc = orc_build_csr(n, m, cs, ce, cw, rowptr, col, J);
    if (rc == 0) {
        orc_problem P = {n, h, rowptr, col, J, sweeps_per_beta, betas, nbetas, order, proposal};
        if (nthreads < 1) nthreads = 1;
        if (nthreads > R) nthreads = R > 0 ? (int)R : 1;
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
        orc_job* jobs = (orc_job*)malloc(sizeof(orc_job) * (size_t)nthreads);
        for (int t = 0; t < nthreads; ++t) {
            orc_job jb = {&P, m, cs, ce, cw, seeds, states, energies, flips, degsum,
                          R * t / nthreads, R * (t + 1) / nthreads};
            jobs[t] = jb;
            if (nthreads == 1) orc_worker(&jobs[t]);
            else pthread_create(&th[t], NULL, orc_worker, &jobs[t]);
        }
        if (nthreads > 1)
            for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
        free(th);
        free(jobs);
    }
    free(rowptr); free(col); free(J);
    return rc;
}

/* Call-level reference semantics (one stream seeded once, continuing across
 * reads in order, cpu_sa.cpp:283-284 and :321-358).  Single threaded. */
int orc_sa_sample_stream(int n, const double* h, int64_t m, const int* cs,
                         const int* ce, const double* cw, int sweeps_per_beta,
                         const double* betas, int nbetas, uint64_t seed, int64_t R,
                         int8_t* states, double* energies, int order, int proposal) {
    int64_t* rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int* col = (int*)malloc(sizeof(int) * (size_t)(2 * m + 1));
    double* J = (double*)malloc(sizeof(double) * (size_t)(2 * m + 1));
    int rc = orc_build_csr(n, m, cs, ce, cw, rowptr, col, J);
    if (rc == 0) {
        orc_problem P = {n, h, rowptr, col, J, sweeps_per_beta, betas, nbetas, order, proposal};
        uint64_t rng[2] = {seed ? seed : UINT64_MAX, 0};
        double* delta = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        for (int64_t r = 0; r < R; ++r) {
            int8_t* s = states + r * n;
            orc_run(&P, s, delta, rng, NULL, NULL);
            energies[r] = orc_state_energy(n, s, h, m, cs, ce, cw);
        }
        free(delta);
    }
    free(rowptr); free(col); free(J);
    return rc;
}
