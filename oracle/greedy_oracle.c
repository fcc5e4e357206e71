/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this; the engine never links it.
 *
 * Plain-C restatement of the reference's steepest-descent solver
 * (dwave/samplers/greedy/src/descent.cpp):
 *   get_flip_energy            descent.cpp:41-62   dE_v = (-2 s_v) * (h_v + sum s_u J)
 *   get_state_energy           descent.cpp:76-95
 *   steepest_gradient_descent_solver (dense)     descent.cpp:117-187
 *       argmin over dE (strictly negative, first index on ties), flip,
 *       negate dE_best, then dE_n -= 4 * s_best * w * s_n for every adjacency
 *       entry of best in insertion order
 *   steepest_gradient_descent  descent.cpp:337-406 (adjacency in coupler
 *       insertion order, index validation, one run per sample)
 *
 * Self-loops are left out of the adjacency (they are constant energy, s^2 = 1,
 * and enter only the final energy): the reference keeps them as neighbours,
 * and its bookkeeping then subtracts 4w from the flipped variable's own flip
 * energy, which for w > 0 can flip that variable back and forth forever.
 * Descents stop after max_steps flips (the reference has no limit).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n;
    const double* h;
    int* rowptr;
    int* col;
    double* J;
    int64_t m;
    const int* cs;
    const int* ce;
    const double* cw;
    int8_t* states;
    double* energies;
    uint32_t* steps;
    uint32_t max_steps;
    int R, nthreads;
} gd_prob;

static double gd_flip_energy(const gd_prob* P, const int8_t* s, int v) {
    double delta = -2 * s[v];
    double contrib = P->h[v];
    for (int k = P->rowptr[v]; k < P->rowptr[v + 1]; ++k) contrib += s[P->col[k]] * P->J[k];
    return delta * contrib;
}

static double gd_state_energy(const gd_prob* P, const int8_t* s) {
    double e = 0.0;
    for (int v = 0; v < P->n; ++v) e += s[v] * P->h[v];
    for (int64_t c = 0; c < P->m; ++c) e += s[P->cs[c]] * P->cw[c] * s[P->ce[c]];
    return e;
}

static uint32_t gd_solve(const gd_prob* P, int8_t* s, double* dE) {
    const int n = P->n;
    if (n < 1) return 0;
    for (int v = 0; v < n; ++v) dE[v] = gd_flip_energy(P, s, v);
    uint32_t steps = 0;
    while (steps < P->max_steps) {
        int best = -1;
        double bestE = 0;
        for (int v = 0; v < n; ++v)
            if (dE[v] < bestE) {
                bestE = dE[v];
                best = v;
            }
        if (best == -1) break;
        s[best] *= -1;
        dE[best] *= -1;
        for (int k = P->rowptr[best]; k < P->rowptr[best + 1]; ++k) {
            const int u = P->col[k];
            dE[u] -= 4 * s[best] * P->J[k] * s[u];
        }
        steps++;
    }
    return steps;
}

static void* gd_worker(void* arg) {
    const gd_prob* P = (const gd_prob*)((void**)arg)[0];
    const int t = (int)(intptr_t)((void**)arg)[1];
    double* dE = (double*)malloc(sizeof(double) * (size_t)(P->n > 0 ? P->n : 1));
    for (int r = t; r < P->R; r += P->nthreads) {
        int8_t* s = P->states + (size_t)r * P->n;
        P->steps[r] = gd_solve(P, s, dE);
        P->energies[r] = gd_state_energy(P, s);
    }
    free(dE);
    return NULL;
}

/* Returns 0, or -2 for a coupler index out of range (descent.cpp:370-372). */
int orc_descent(int n, const double* h, int64_t m, const int* cs, const int* ce, const double* cw, int R,
                int8_t* states, double* energies, uint32_t* steps, uint32_t max_steps, int nthreads) {
    for (int64_t c = 0; c < m; ++c)
        if (cs[c] < 0 || ce[c] < 0 || cs[c] >= n || ce[c] >= n) return -2;
    gd_prob P;
    memset(&P, 0, sizeof(P));
    P.n = n;
    P.h = h;
    P.m = m;
    P.cs = cs;
    P.ce = ce;
    P.cw = cw;
    P.states = states;
    P.energies = energies;
    P.steps = steps;
    P.max_steps = max_steps;
    P.R = R;
    P.rowptr = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int64_t c = 0; c < m; ++c) {
        if (cs[c] == ce[c]) continue;
        P.rowptr[cs[c] + 1]++;
        P.rowptr[ce[c] + 1]++;
    }
    for (int v = 0; v < n; ++v) P.rowptr[v + 1] += P.rowptr[v];
    P.col = (int*)malloc(sizeof(int) * (size_t)(2 * m > 0 ? 2 * m : 1));
    P.J = (double*)malloc(sizeof(double) * (size_t)(2 * m > 0 ? 2 * m : 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    memcpy(fill, P.rowptr, sizeof(int) * (size_t)(n + 1));
    for (int64_t c = 0; c < m; ++c) {
        if (cs[c] == ce[c]) continue;
        int k = fill[cs[c]]++;
        P.col[k] = ce[c];
        P.J[k] = cw[c];
        k = fill[ce[c]]++;
        P.col[k] = cs[c];
        P.J[k] = cw[c];
    }
    free(fill);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R > 0 ? R : 1;
    P.nthreads = nthreads;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    void** args = (void**)malloc(sizeof(void*) * 2 * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        args[2 * t] = &P;
        args[2 * t + 1] = (void*)(intptr_t)t;
        pthread_create(&th[t], NULL, gd_worker, &args[2 * t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(args);
    free(P.rowptr);
    free(P.col);
    free(P.J);
    return 0;
}
