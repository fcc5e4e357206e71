/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this; the engine never links it.
 *
 * CPU restatement of the engine's discrete-imaginary-time path-integral Monte
 * Carlo (the MI355X counterpart of the reference's localPIMC, whose algorithm
 * this follows in discrete form):
 *
 *   reference (continuous time, 2^16 resolution)     this restatement (P slices)
 *   ------------------------------------------------ ---------------------------------
 *   Poisson break proposals at rate Hd               cut each equal-spin bond between
 *     (localPIMC.cpp:527-538)                          slices k-1,k with prob tanh(Hd/P)
 *   existing breaks are forced interfaces            kinks are forced cuts
 *     (localPIMC.cpp:354-361)
 *   per-domain field, sum of h and J*s over the      per-segment field F = sum over the
 *     domain's share of imaginary time                segment's slices of h_i + sum_j J_ij s_jk
 *     (localPIMC.cpp:279-348)
 *   domain spin by heat bath, P(s) ~ exp(-Hp*F*s)    segment spin by heat bath,
 *     (GibbsSamplePM1, localPIMC.cpp:551-556)         P(s) ~ exp(-(Hp/P)*F*s)
 *   breaks where adjacent domains differ (:374-382) kinks where adjacent slices differ
 *   projected state = first-slice spin (:181-185)    projected state = slice 0
 *
 * This is the Swendsen-Wang update along imaginary time of the Trotter model
 *     Z_P = Tr[(exp(-(Hp/P) Hz) exp((Hd/P) sum_i sigma^x_i))^P],
 *     Hz  = sum_i h_i sigma^z_i + sum_(i,j) J_ij sigma^z_i sigma^z_j,
 * whose equilibrium the tests pin against the exact transfer matrix.
 *
 * Order, draws and arithmetic are those of the HIP kernel, so that the two
 * agree bit for bit:
 *   - sites are visited colour class by colour class (greedy colouring in
 *     label order of the coupler graph without self-loops); sites of one
 *     class are never adjacent, so the order inside a class is immaterial;
 *   - every (read, sweep, site) update draws from its own xoshiro128+
 *     stream whose state is hashed from the read seed, the global sweep
 *     index and the site;
 *   - the heat-bath exponent y = (2 Hp / P) log2(e) F is formed in f32 as
 *     len * A_i + S * B when all |J| are equal (S the integer count of
 *     satisfied neighbour slices), else from the f64 sum F, and clamped to
 *     [-92, 92];
 *   - exp2 / log2 are evaluated by fixed f32 polynomials with fmaf, never by
 *     libm, so the CPU and the GPU round identically;
 *   - self-loops are constant energy (s_i^2 = 1): they enter the energy but
 *     not the dynamics.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PO_NONE 0
#define PO_PROP 1
#define PO_ALL 2

#define PO_KIND_TABLE 0
#define PO_KIND_SIGNED 1

static uint64_t po_mix64(uint64_t z) { /* splitmix64 finalizer */
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebull;
    z ^= z >> 31;
    return z;
}

static uint32_t po_fmix32(uint32_t h) { /* murmur3 finalizer */
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

static inline uint32_t po_rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

typedef struct {
    uint32_t s0, s1, s2, s3;
} po_rng;

static inline uint32_t po_next(po_rng* g) { /* xoshiro128+ */
    const uint32_t r = g->s0 + g->s3;
    const uint32_t t = g->s1 << 9;
    g->s2 ^= g->s0;
    g->s3 ^= g->s1;
    g->s1 ^= g->s2;
    g->s0 ^= g->s3;
    g->s2 ^= t;
    g->s3 = po_rotl(g->s3, 11);
    return r;
}

static inline uint64_t po_sweep_key(uint64_t seed, uint64_t g) {
    return po_mix64(seed + (g + 1) * 0x9E3779B97F4A7C15ull);
}

/* Two murmur3-finalised words of (key, site); the other two state words are
 * fixed xors of them, so the state is never all zero. */
static inline po_rng po_task_rng(uint64_t key, uint32_t site) {
    po_rng g;
    g.s0 = po_fmix32((uint32_t)key + site);
    g.s1 = po_fmix32((uint32_t)(key >> 32) ^ (0x9E3779B9u * (site + 1u)));
    g.s2 = g.s0 ^ 0x9E3779B9u;
    g.s3 = g.s1 ^ 0x7F4A7C15u;
    return g;
}

static inline float po_asf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t po_asu(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* 2^y for y in [-93, 93]: 2^frac by a degree-6 polynomial (rel. err. 6e-8). */
static float po_exp2f(float y) {
    float fi = floorf(y);
    float f = y - fi;
    float p = 0x1.c6e292p-13f;
    p = fmaf(p, f, 0x1.46301cp-10f);
    p = fmaf(p, f, 0x1.3d24eap-7f);
    p = fmaf(p, f, 0x1.c68562p-5f);
    p = fmaf(p, f, 0x1.ebfd9ap-3f);
    p = fmaf(p, f, 0x1.62e42ap-1f);
    p = fmaf(p, f, 0x1p+0f);
    return po_asf((uint32_t)((int32_t)po_asu(p) + (int32_t)fi * (1 << 23)));
}

/* log2 of U = v * 2^-24, v in [1, 2^24] (an exact float): exponent plus a
 * degree-7 polynomial in t = m - 1, m in [sqrt(1/2), sqrt(2)) (err. 3e-7). */
static float po_log2u(uint32_t v) {
    float fv = (float)v;
    uint32_t bits = po_asu(fv);
    int e = (int)(bits >> 23) - 127;
    float mm = po_asf((bits & 0x7fffffu) | 0x3f800000u);
    if (mm > 0x1.6a09e6p+0f) {
        mm = mm * 0.5f;
        e += 1;
    }
    float t = mm - 1.0f;
    float q = -0x1.239648p-3f;
    q = fmaf(q, t, 0x1.dc1d94p-3f);
    q = fmaf(q, t, -0x1.fe9c4ap-3f);
    q = fmaf(q, t, 0x1.26355p-2f);
    q = fmaf(q, t, -0x1.70de64p-2f);
    q = fmaf(q, t, 0x1.ec7536p-2f);
    q = fmaf(q, t, -0x1.71552cp-1f);
    q = fmaf(q, t, 0x1.715476p+0f);
    float tq = t * q;
    return (float)(e - 24) + tq;
}

/* Geometric gap (number of kept bonds before the next cut). */
static inline int po_gap(uint32_t u, float ginv) {
    float y = po_log2u((u >> 8) + 1u) * ginv;
    y = fminf(y, 64.0f);
    return (int)y;
}

/* Per schedule step: slice inverse temperature Hp/P and the cut law. */
void orc_pimc_step_consts(double hp, double hd, int P, double* bs, float* ginv, int* mode) {
    *bs = hp / (double)P;
    *ginv = 0.0f;
    if (!(hd > 0.0)) {
        *mode = PO_NONE;
        return;
    }
    double p = tanh(hd / (double)P);
    if (p >= 1.0) {
        *mode = PO_ALL;
        return;
    }
    float g = (float)(0x1.62e42fefa39efp-1 / log1p(-p)); /* ln 2 / ln(1 - p) */
    if (!isfinite(g)) { /* p below ~1e-38: no cut is ever proposed */
        *mode = PO_NONE;
        return;
    }
    *mode = PO_PROP;
    *ginv = g;
}

typedef struct {
    int n, P, kind, nclass;
    uint32_t maskP;
    double J0;
    const double* h;
    int* rowptr;
    int* col;
    double* J;
    uint32_t* neg; /* SIGNED: maskP where J < 0 */
    int* class_off;
    int* class_sites;
    /* run */
    int spb, nsteps, interval, ncol;
    const double* bs;
    const float* ginv;
    const int* mode;
    const uint64_t* seeds;
    uint32_t* W;
    int8_t* stats;
    uint64_t* flips;
    int R, nthreads;
} po_prob;

/* Heat-bath exponent y = bs2l * F for the segment m (len slices) of site i,
 * F = sum over the segment's slices of h_i + sum_j J_ij s_jk. */
static float po_y(const po_prob* pr, const uint32_t* W, int i, uint32_t m, int len, double bs2l, float A,
                  float B) {
    const int b = pr->rowptr[i], e = pr->rowptr[i + 1];
    float y;
    if (pr->kind == PO_KIND_SIGNED) {
        int S = 0;
        for (int k = b; k < e; ++k) S += __builtin_popcount((W[pr->col[k]] ^ pr->neg[k]) & m);
        const float t = (float)S * B;
        y = fmaf((float)len, A, t);
    } else {
        double F = (double)len * pr->h[i];
        for (int k = b; k < e; ++k) F = F + pr->J[k] * (double)(2 * __builtin_popcount(W[pr->col[k]] & m) - len);
        y = (float)(F * bs2l);
    }
    return fminf(fmaxf(y, -92.0f), 92.0f);
}

static uint32_t po_update(const po_prob* pr, uint32_t* W, int i, po_rng* g, double bs2l, float ginv, int mode,
                          float B) {
    const int P = pr->P;
    const uint32_t maskP = pr->maskP;
    const uint32_t w0 = W[i];
    uint32_t w = w0;
    float A = 0.0f;
    if (pr->kind == PO_KIND_SIGNED) {
        const int deg = pr->rowptr[i + 1] - pr->rowptr[i];
        A = (float)((pr->h[i] - pr->J0 * (double)deg) * bs2l);
    }
    uint32_t rot = ((w << 1) | (w >> (P - 1))) & maskP;
    uint32_t cuts = w ^ rot;
    if (mode == PO_ALL) {
        cuts = maskP;
    } else if (mode == PO_PROP) {
        int pos = po_gap(po_next(g), ginv);
        while (pos < P) {
            cuts |= 1u << pos;
            pos += 1 + po_gap(po_next(g), ginv);
        }
    }
    if (__builtin_popcount(cuts) <= 1) cuts = 1u;
    const int first = __builtin_ctz(cuts);
    uint32_t rem = cuts & (cuts - 1u);
    int a = first;
    for (;;) {
        int b;
        if (rem) {
            b = __builtin_ctz(rem);
            rem &= rem - 1u;
        } else {
            b = first + P;
        }
        const int len = b - a;
        const uint64_t m64 = (1ull << b) - (1ull << a);
        const uint32_t m = (uint32_t)((m64 | (m64 >> P)) & maskP);
        const float y = po_y(pr, W, i, m, len, bs2l, A, B);
        const float e = po_exp2f(y);
        const float u = po_asf(0x3F800000u | (po_next(g) >> 9)) - 1.0f;
        const float t = 1.0f + e;
        const int plus = (u * t) < 1.0f;
        w = (w & ~m) | (plus ? m : 0u);
        if (b == first + P) break;
        a = b;
    }
    W[i] = w;
    return (uint32_t)__builtin_popcount(w ^ w0);
}

static void po_run_read(const po_prob* pr, int r) {
    const int n = pr->n;
    uint32_t* W = pr->W + (size_t)r * n;
    const uint64_t seed = pr->seeds[r];
    uint64_t nflip = 0;
    for (int k = 0; k < pr->nsteps; ++k) {
        const double bs2l = 2.0 * pr->bs[k] * 0x1.71547652b82fep+0; /* 2 (Hp/P) log2(e) */
        const float B = (float)(2.0 * pr->J0 * bs2l);
        for (int sw = 0; sw < pr->spb; ++sw) {
            const uint64_t gidx = (uint64_t)k * pr->spb + sw;
            const uint64_t key = po_sweep_key(seed, gidx);
            for (int c = 0; c < pr->nclass; ++c)
                for (int q = pr->class_off[c]; q < pr->class_off[c + 1]; ++q) {
                    const int i = pr->class_sites[q];
                    po_rng g = po_task_rng(key, (uint32_t)i);
                    nflip += po_update(pr, W, i, &g, bs2l, pr->ginv[k], pr->mode[k], B);
                }
        }
        if (pr->stats && pr->interval > 0 && k % pr->interval == 0) {
            int8_t* row = pr->stats + ((size_t)r * pr->ncol + k / pr->interval) * n;
            for (int i = 0; i < n; ++i) row[i] = (W[i] & 1u) ? 1 : -1;
        }
    }
    if (pr->flips) pr->flips[r] = nflip;
}

typedef struct {
    const po_prob* pr;
    int t;
} po_job;

static void* po_worker(void* arg) {
    const po_job* j = (const po_job*)arg;
    const po_prob* pr = j->pr;
    for (int r = j->t; r < pr->R; r += pr->nthreads) po_run_read(pr, r);
    return NULL;
}

/* Greedy colouring in label order; returns the number of classes and fills
 * class_off / class_sites (ascending labels inside a class). */
static int po_colour(int n, const int* rowptr, const int* col, int* class_off, int* class_sites) {
    int* colour = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int* mark = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    int ncol = 0;
    for (int c = 0; c <= n; ++c) mark[c] = -1;
    for (int v = 0; v < n; ++v) {
        for (int k = rowptr[v]; k < rowptr[v + 1]; ++k) {
            const int u = col[k];
            if (u < v) mark[colour[u]] = v;
        }
        int c = 0;
        while (mark[c] == v) ++c;
        colour[v] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    for (int c = 0; c <= ncol; ++c) class_off[c] = 0;
    for (int v = 0; v < n; ++v) class_off[colour[v] + 1]++;
    for (int c = 0; c < ncol; ++c) class_off[c + 1] += class_off[c];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(ncol + 1));
    memcpy(fill, class_off, sizeof(int) * (size_t)(ncol + 1));
    for (int v = 0; v < n; ++v) class_sites[fill[colour[v]]++] = v;
    free(fill);
    free(mark);
    free(colour);
    return ncol;
}

/* Number of colour classes and the class of every site (for tests). */
int orc_pimc_colouring(int n, int64_t m, const int* cs, const int* ce, int* colour_out) {
    int* deg = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int64_t c = 0; c < m; ++c) {
        if (cs[c] < 0 || ce[c] < 0 || cs[c] >= n || ce[c] >= n) { free(deg); return -2; }
        if (cs[c] == ce[c]) continue;
        deg[cs[c] + 1]++;
        deg[ce[c] + 1]++;
    }
    for (int v = 0; v < n; ++v) deg[v + 1] += deg[v];
    int* col = (int*)malloc(sizeof(int) * (size_t)(deg[n] > 0 ? deg[n] : 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    memcpy(fill, deg, sizeof(int) * (size_t)(n + 1));
    for (int64_t c = 0; c < m; ++c) {
        if (cs[c] == ce[c]) continue;
        col[fill[cs[c]]++] = ce[c];
        col[fill[ce[c]]++] = cs[c];
    }
    int* off = (int*)malloc(sizeof(int) * (size_t)(n + 2));
    int* sites = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int nc = po_colour(n, deg, col, off, sites);
    for (int c = 0; c < nc; ++c)
        for (int q = off[c]; q < off[c + 1]; ++q) colour_out[sites[q]] = c;
    free(off); free(sites); free(fill); free(col); free(deg);
    return nc;
}

/*
 * Anneal R reads.  W (R x n, in/out) holds one bit per slice (bit k = slice k,
 * set = spin +1).  stats, when interval > 0, is (R, 1 + (nsteps-1)/interval, n)
 * int8 and receives slice 0 after every interval-th schedule step
 * (localPIMC.cpp:140-143).  flips (R, may be NULL) receives the number of
 * slice spins each read changed.  colour (n, may be NULL) gives the class of
 * every site (any proper colouring: classes visited in increasing colour);
 * NULL uses the label-order greedy colouring.  Returns 0, -1 for bad
 * arguments or an improper colouring, -2 for a coupler index out of range
 * (localPIMC.cpp:660-662).
 */
int orc_pimc_sample(int n, const double* h, int64_t m, const int* cs, const int* ce, const double* cw, int R,
                    int spb, const double* hp, const double* hd, int nsteps, int P, const uint64_t* seeds,
                    uint32_t* W, int8_t* stats, int interval, uint64_t* flips, const int* colour,
                    int nthreads) {
    if (n < 0 || R < 0 || spb < 0 || nsteps < 0 || P < 1 || P > 32) return -1;
    for (int64_t c = 0; c < m; ++c)
        if (cs[c] < 0 || ce[c] < 0 || cs[c] >= n || ce[c] >= n) return -2;
    if ((int64_t)n * R == 0) return 0;
    po_prob pr;
    memset(&pr, 0, sizeof(pr));
    pr.n = n;
    pr.P = P;
    pr.maskP = P == 32 ? 0xffffffffu : ((1u << P) - 1u);
    pr.h = h;
    /* CSR without self-loops, coupler order (localPIMC.cpp:652-674). */
    pr.rowptr = (int*)calloc((size_t)n + 1, sizeof(int));
    int64_t m2 = 0;
    double J0 = -1.0;
    int same = 1;
    for (int64_t c = 0; c < m; ++c) {
        if (cs[c] == ce[c]) continue;
        pr.rowptr[cs[c] + 1]++;
        pr.rowptr[ce[c] + 1]++;
        m2 += 2;
        const double a = fabs(cw[c]);
        if (J0 < 0) J0 = a;
        else if (a != J0) same = 0;
    }
    for (int v = 0; v < n; ++v) pr.rowptr[v + 1] += pr.rowptr[v];
    pr.kind = (same && J0 > 0 && n <= 32768) ? PO_KIND_SIGNED : PO_KIND_TABLE;
    pr.J0 = J0 > 0 ? J0 : 0.0;
    pr.col = (int*)malloc(sizeof(int) * (size_t)(m2 > 0 ? m2 : 1));
    pr.J = (double*)malloc(sizeof(double) * (size_t)(m2 > 0 ? m2 : 1));
    pr.neg = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(m2 > 0 ? m2 : 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    memcpy(fill, pr.rowptr, sizeof(int) * (size_t)(n + 1));
    for (int64_t c = 0; c < m; ++c) {
        const int u = cs[c], v = ce[c];
        if (u == v) continue;
        int k = fill[u]++;
        pr.col[k] = v;
        pr.J[k] = cw[c];
        pr.neg[k] = cw[c] < 0 ? pr.maskP : 0u;
        k = fill[v]++;
        pr.col[k] = u;
        pr.J[k] = cw[c];
        pr.neg[k] = cw[c] < 0 ? pr.maskP : 0u;
    }
    free(fill);
    pr.class_off = (int*)malloc(sizeof(int) * (size_t)(n + 2));
    pr.class_sites = (int*)malloc(sizeof(int) * (size_t)n);
    if (colour) {
        int nc = 0, bad = 0;
        for (int v = 0; v < n; ++v) {
            if (colour[v] < 0 || colour[v] >= n) bad = 1;
            else if (colour[v] + 1 > nc) nc = colour[v] + 1;
            for (int k = pr.rowptr[v]; k < pr.rowptr[v + 1]; ++k)
                if (colour[pr.col[k]] == colour[v]) bad = 1;
        }
        if (bad) {
            free(pr.rowptr); free(pr.col); free(pr.J); free(pr.neg); free(pr.class_off); free(pr.class_sites);
            return -1;
        }
        for (int c = 0; c <= nc; ++c) pr.class_off[c] = 0;
        for (int v = 0; v < n; ++v) pr.class_off[colour[v] + 1]++;
        for (int c = 0; c < nc; ++c) pr.class_off[c + 1] += pr.class_off[c];
        int* cf = (int*)malloc(sizeof(int) * (size_t)(nc + 1));
        memcpy(cf, pr.class_off, sizeof(int) * (size_t)(nc + 1));
        for (int v = 0; v < n; ++v) pr.class_sites[cf[colour[v]]++] = v;
        free(cf);
        pr.nclass = nc;
    } else {
        pr.nclass = po_colour(n, pr.rowptr, pr.col, pr.class_off, pr.class_sites);
    }
    double* bs = (double*)malloc(sizeof(double) * (size_t)(nsteps > 0 ? nsteps : 1));
    float* gi = (float*)malloc(sizeof(float) * (size_t)(nsteps > 0 ? nsteps : 1));
    int* md = (int*)malloc(sizeof(int) * (size_t)(nsteps > 0 ? nsteps : 1));
    for (int k = 0; k < nsteps; ++k) orc_pimc_step_consts(hp[k], hd[k], P, &bs[k], &gi[k], &md[k]);
    pr.spb = spb;
    pr.nsteps = nsteps;
    pr.interval = interval;
    pr.ncol = interval > 0 ? 1 + (nsteps - 1) / interval : 0;
    if (nsteps == 0) pr.ncol = interval > 0 ? 1 : 0;
    pr.bs = bs;
    pr.ginv = gi;
    pr.mode = md;
    pr.seeds = seeds;
    pr.W = W;
    pr.stats = stats;
    pr.flips = flips;
    pr.R = R;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R;
    pr.nthreads = nthreads;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    po_job* jobs = (po_job*)malloc(sizeof(po_job) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].pr = &pr;
        jobs[t].t = t;
        pthread_create(&th[t], NULL, po_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(jobs); free(bs); free(gi); free(md);
    free(pr.rowptr); free(pr.col); free(pr.J); free(pr.neg); free(pr.class_off); free(pr.class_sites);
    return 0;
}

/*
 * Chain updates (qubits_per_update > 1): the discrete form of the reference's
 * chainUpdate (localPIMC.cpp:385-452, makeDomainGraph :462-525,
 * depthFirstDomainAssignment :454-460, pNotJoin :558).  Chain c holds qubits
 * c*Q .. c*Q+Q-1 (c < n / Q; later qubits are never updated, as in
 * localPIMC.cpp:122,146).  One chain update:
 *   1. every qubit of the chain gets its cuts as a site update does (kinks
 *      plus proposed cuts, qubit by qubit); its segments are the nodes;
 *   2. for consecutive qubits q, q+1 the pieces between consecutive cuts of
 *      either are visited in slice order, from the first cut; where the two
 *      spins agree the segments holding the piece are joined with
 *      probability 1 - exp(2 Jchain (Hp/P) len) (the reference joins when
 *      pNotJoin(len) < u; Jchain >= 0 never joins).  Splitting a piece
 *      leaves the law unchanged: exp is multiplicative in len;
 *   3. node field F = sum over the segment's slices of h_i + sum_j J_ij s_jk
 *      with every neighbour's worldline as it was before the update (f64),
 *      added per component in node order (union by smaller node index);
 *   4. each component, in order of its smallest node, takes its spin by heat
 *      bath P(s) ~ exp(-(Hp/P) F s) as a segment does.
 * A sweep makes Q passes over the chains, class by class (greedy colouring in
 * label order of the chain graph: chains are adjacent when a coupler joins
 * them), as the reference makes num_vars random chain updates per sweep.  The
 * stream of a (read, sweep, pass, chain) update is po_task_rng(key, chain)
 * with key = po_sweep_key(seed, (step * spb + sweep) * Q + pass).
 */
typedef struct {
    po_prob base;
    int Q, nchains;
    double jchain;
    int* chain_off; /* classes of chains */
    int* chain_list;
} po_chains;

static int po_find(uint16_t* parent, int x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

/* Node (segment) of qubit q holding slice a: the last cut at or below a, else
 * the last segment (it wraps past slice P-1). */
static inline int po_seg_of(uint32_t cuts, int a) {
    const uint32_t upto = a >= 31 ? 0xffffffffu : ((2u << a) - 1u);
    const int r = __builtin_popcount(cuts & upto) - 1;
    return r >= 0 ? r : __builtin_popcount(cuts) - 1;
}

static uint32_t po_chain_update(const po_chains* ch, uint32_t* W, int c, po_rng* g, double bs2l, float ginv,
                                int mode) {
    const po_prob* pr = &ch->base;
    const int P = pr->P, Q = ch->Q, q0 = c * Q;
    const uint32_t maskP = pr->maskP;
    uint32_t* cuts = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)Q);
    int N = 0;
    /* 1. cuts, qubit by qubit */
    for (int q = 0; q < Q; ++q) {
        const uint32_t w = W[q0 + q];
        uint32_t cu = w ^ (((w << 1) | (w >> (P - 1))) & maskP);
        if (mode == PO_ALL) {
            cu = maskP;
        } else if (mode == PO_PROP) {
            int pos = po_gap(po_next(g), ginv);
            while (pos < P) {
                cu |= 1u << pos;
                pos += 1 + po_gap(po_next(g), ginv);
            }
        }
        if (__builtin_popcount(cu) <= 1) cu = 1u;
        cuts[q] = cu;
        N += __builtin_popcount(cu);
    }
    uint16_t* parent = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)N);
    double* Fc = (double*)malloc(sizeof(double) * (size_t)N);
    uint8_t* up = (uint8_t*)calloc((size_t)N, 1);
    for (int x = 0; x < N; ++x) parent[x] = (uint16_t)x;
    /* 2. joins between consecutive qubits, piece by piece of their joint cuts */
    int bq = 0;
    for (int q = 0; q + 1 < Q; ++q) {
        const int bq1 = bq + __builtin_popcount(cuts[q]);
        const uint32_t U = cuts[q] | cuts[q + 1];
        const uint32_t wa = W[q0 + q], wb = W[q0 + q + 1];
        const int first = __builtin_ctz(U);
        uint32_t rem = U & (U - 1u);
        int a = first;
        for (;;) {
            const int b = rem ? __builtin_ctz(rem) : first + P;
            if (rem) rem &= rem - 1u;
            if (((wa ^ wb) >> a & 1u) == 0u) {
                float y = (float)(ch->jchain * (double)(b - a) * bs2l);
                y = fminf(fmaxf(y, -92.0f), 92.0f);
                const float pnot = po_exp2f(y);
                const float u = po_asf(0x3F800000u | (po_next(g) >> 9)) - 1.0f;
                if (pnot < u) {
                    const int x = po_find(parent, bq + po_seg_of(cuts[q], a));
                    const int z = po_find(parent, bq1 + po_seg_of(cuts[q + 1], a));
                    if (x != z) parent[x > z ? x : z] = (uint16_t)(x < z ? x : z);
                }
            }
            if (b == first + P) break;
            a = b;
        }
        bq = bq1;
    }
    /* 3. node fields (neighbour worldlines as before the update), summed per
     * component in node order; a component's root is its smallest node */
    int x = 0;
    for (int q = 0; q < Q; ++q) {
        const int i = q0 + q;
        const uint32_t cu = cuts[q];
        const int first = __builtin_ctz(cu);
        uint32_t rem = cu & (cu - 1u);
        int a = first;
        for (;;) {
            const int b = rem ? __builtin_ctz(rem) : first + P;
            if (rem) rem &= rem - 1u;
            const int len = b - a;
            const uint64_t m64 = (1ull << b) - (1ull << a);
            const uint32_t m = (uint32_t)((m64 | (m64 >> P)) & maskP);
            double f = (double)len * pr->h[i];
            for (int k = pr->rowptr[i]; k < pr->rowptr[i + 1]; ++k)
                f = f + pr->J[k] * (double)(2 * __builtin_popcount(W[pr->col[k]] & m) - len);
            const int r = po_find(parent, x);
            Fc[r] = r == x ? f : Fc[r] + f;
            ++x;
            if (b == first + P) break;
            a = b;
        }
    }
    /* 4. one heat-bath spin per component, in order of the roots */
    for (x = 0; x < N; ++x) {
        if (parent[x] != x) continue;
        float y = (float)(Fc[x] * bs2l);
        y = fminf(fmaxf(y, -92.0f), 92.0f);
        const float e = po_exp2f(y);
        const float u = po_asf(0x3F800000u | (po_next(g) >> 9)) - 1.0f;
        up[x] = (u * (1.0f + e)) < 1.0f;
    }
    /* 5. new worldlines */
    uint32_t changed = 0;
    x = 0;
    for (int q = 0; q < Q; ++q) {
        const uint32_t cu = cuts[q];
        const int first = __builtin_ctz(cu);
        uint32_t rem = cu & (cu - 1u);
        int a = first;
        uint32_t w = 0;
        for (;;) {
            const int b = rem ? __builtin_ctz(rem) : first + P;
            if (rem) rem &= rem - 1u;
            const uint64_t m64 = (1ull << b) - (1ull << a);
            const uint32_t m = (uint32_t)((m64 | (m64 >> P)) & maskP);
            if (up[po_find(parent, x)]) w |= m;
            ++x;
            if (b == first + P) break;
            a = b;
        }
        changed += (uint32_t)__builtin_popcount(w ^ W[q0 + q]);
        W[q0 + q] = w;
    }
    free(up); free(Fc); free(parent); free(cuts);
    return changed;
}

static void po_chain_run_read(const po_chains* ch, int r) {
    const po_prob* pr = &ch->base;
    const int n = pr->n;
    uint32_t* W = pr->W + (size_t)r * n;
    const uint64_t seed = pr->seeds[r];
    uint64_t nflip = 0;
    for (int k = 0; k < pr->nsteps; ++k) {
        const double bs2l = 2.0 * pr->bs[k] * 0x1.71547652b82fep+0;
        for (int sw = 0; sw < pr->spb; ++sw)
            for (int pass = 0; pass < ch->Q; ++pass) {
                const uint64_t gidx = ((uint64_t)k * pr->spb + sw) * (uint64_t)ch->Q + pass;
                const uint64_t key = po_sweep_key(seed, gidx);
                for (int cl = 0; cl < pr->nclass; ++cl)
                    for (int t = ch->chain_off[cl]; t < ch->chain_off[cl + 1]; ++t) {
                        const int c = ch->chain_list[t];
                        po_rng g = po_task_rng(key, (uint32_t)c);
                        nflip += po_chain_update(ch, W, c, &g, bs2l, pr->ginv[k], pr->mode[k]);
                    }
            }
        if (pr->stats && pr->interval > 0 && k % pr->interval == 0) {
            int8_t* row = pr->stats + ((size_t)r * pr->ncol + k / pr->interval) * n;
            for (int i = 0; i < n; ++i) row[i] = (W[i] & 1u) ? 1 : -1;
        }
    }
    if (pr->flips) pr->flips[r] = nflip;
}

typedef struct {
    const po_chains* ch;
    int t;
} po_chain_job;

static void* po_chain_worker(void* arg) {
    const po_chain_job* j = (const po_chain_job*)arg;
    const po_prob* pr = &j->ch->base;
    for (int r = j->t; r < pr->R; r += pr->nthreads) po_chain_run_read(j->ch, r);
    return NULL;
}

/* Greedy colouring in label order of the chain graph (chains c != c' adjacent
 * when a coupler joins a qubit of c to one of c'); returns the number of
 * classes, fills off (nclass + 1) and list (nchains) class by class. */
static int po_chain_colour(int n, int Q, const int* rowptr, const int* col, int* off, int* list) {
    const int nch = n / Q;
    int* colour = (int*)malloc(sizeof(int) * (size_t)(nch > 0 ? nch : 1));
    int* mark = (int*)malloc(sizeof(int) * (size_t)(nch + 1));
    int nc = 0;
    for (int c = 0; c <= nch; ++c) mark[c] = -1;
    for (int c = 0; c < nch; ++c) {
        for (int i = c * Q; i < c * Q + Q; ++i)
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
                const int d = col[k] / Q;
                if (d < c) mark[colour[d]] = c;
            }
        int x = 0;
        while (mark[x] == c) ++x;
        colour[c] = x;
        if (x + 1 > nc) nc = x + 1;
    }
    for (int x = 0; x <= nc; ++x) off[x] = 0;
    for (int c = 0; c < nch; ++c) off[colour[c] + 1]++;
    for (int x = 0; x < nc; ++x) off[x + 1] += off[x];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(nc + 1));
    memcpy(fill, off, sizeof(int) * (size_t)(nc + 1));
    for (int c = 0; c < nch; ++c) list[fill[colour[c]]++] = c;
    free(fill); free(mark); free(colour);
    return nc;
}

/*
 * Anneal R reads with chain updates of Q >= 1 qubits (arguments as
 * orc_pimc_sample, plus Q and the chain coupling jchain).  chain_colour
 * (n / Q entries, may be NULL) gives the class of every chain (any colouring
 * in which chains of one class share no coupler; NULL: label-order greedy).
 * Returns 0, -1 for bad arguments or an improper colouring, -2 for a coupler
 * index out of range.
 */
int orc_pimc_sample_chains(int n, const double* h, int64_t m, const int* cs, const int* ce, const double* cw,
                           int R, int spb, const double* hp, const double* hd, int nsteps, int P,
                           const uint64_t* seeds, uint32_t* W, int8_t* stats, int interval, uint64_t* flips,
                           int Q, double jchain, const int* chain_colour, int nthreads) {
    if (n < 0 || R < 0 || spb < 0 || nsteps < 0 || P < 1 || P > 32 || Q < 1) return -1;
    for (int64_t c = 0; c < m; ++c)
        if (cs[c] < 0 || ce[c] < 0 || cs[c] >= n || ce[c] >= n) return -2;
    if ((int64_t)n * R == 0) return 0;
    po_chains ch;
    memset(&ch, 0, sizeof(ch));
    po_prob* pr = &ch.base;
    pr->n = n;
    pr->P = P;
    pr->maskP = P == 32 ? 0xffffffffu : ((1u << P) - 1u);
    pr->h = h;
    pr->rowptr = (int*)calloc((size_t)n + 1, sizeof(int));
    int64_t m2 = 0;
    for (int64_t c = 0; c < m; ++c) {
        if (cs[c] == ce[c]) continue;
        pr->rowptr[cs[c] + 1]++;
        pr->rowptr[ce[c] + 1]++;
        m2 += 2;
    }
    for (int v = 0; v < n; ++v) pr->rowptr[v + 1] += pr->rowptr[v];
    pr->kind = PO_KIND_TABLE;
    pr->col = (int*)malloc(sizeof(int) * (size_t)(m2 > 0 ? m2 : 1));
    pr->J = (double*)malloc(sizeof(double) * (size_t)(m2 > 0 ? m2 : 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    memcpy(fill, pr->rowptr, sizeof(int) * (size_t)(n + 1));
    for (int64_t c = 0; c < m; ++c) {
        const int u = cs[c], v = ce[c];
        if (u == v) continue;
        int k = fill[u]++;
        pr->col[k] = v;
        pr->J[k] = cw[c];
        k = fill[v]++;
        pr->col[k] = u;
        pr->J[k] = cw[c];
    }
    free(fill);
    ch.Q = Q;
    ch.jchain = jchain;
    ch.nchains = n / Q;
    ch.chain_off = (int*)malloc(sizeof(int) * (size_t)(ch.nchains + 2));
    ch.chain_list = (int*)malloc(sizeof(int) * (size_t)(ch.nchains > 0 ? ch.nchains : 1));
    if (chain_colour) {
        const int nch = ch.nchains;
        int nc = 0, bad = 0;
        for (int c = 0; c < nch; ++c) {
            if (chain_colour[c] < 0 || chain_colour[c] >= nch) {
                bad = 1;
                continue;
            }
            if (chain_colour[c] + 1 > nc) nc = chain_colour[c] + 1;
        }
        for (int c = 0; c < nch && !bad; ++c)
            for (int i = c * Q; i < c * Q + Q; ++i)
                for (int k = pr->rowptr[i]; k < pr->rowptr[i + 1]; ++k) {
                    const int d = pr->col[k] / Q;
                    if (d != c && d < nch && chain_colour[d] == chain_colour[c]) bad = 1;
                }
        if (bad) {
            free(pr->rowptr); free(pr->col); free(pr->J); free(ch.chain_off); free(ch.chain_list);
            return -1;
        }
        for (int x = 0; x <= nc; ++x) ch.chain_off[x] = 0;
        for (int c = 0; c < nch; ++c) ch.chain_off[chain_colour[c] + 1]++;
        for (int x = 0; x < nc; ++x) ch.chain_off[x + 1] += ch.chain_off[x];
        int* cf = (int*)malloc(sizeof(int) * (size_t)(nc + 1));
        memcpy(cf, ch.chain_off, sizeof(int) * (size_t)(nc + 1));
        for (int c = 0; c < nch; ++c) ch.chain_list[cf[chain_colour[c]]++] = c;
        free(cf);
        pr->nclass = nc;
    } else {
        pr->nclass = po_chain_colour(n, Q, pr->rowptr, pr->col, ch.chain_off, ch.chain_list);
    }
    double* bs = (double*)malloc(sizeof(double) * (size_t)(nsteps > 0 ? nsteps : 1));
    float* gi = (float*)malloc(sizeof(float) * (size_t)(nsteps > 0 ? nsteps : 1));
    int* md = (int*)malloc(sizeof(int) * (size_t)(nsteps > 0 ? nsteps : 1));
    for (int k = 0; k < nsteps; ++k) orc_pimc_step_consts(hp[k], hd[k], P, &bs[k], &gi[k], &md[k]);
    pr->spb = spb;
    pr->nsteps = nsteps;
    pr->interval = interval;
    pr->ncol = interval > 0 ? (nsteps > 0 ? 1 + (nsteps - 1) / interval : 1) : 0;
    pr->bs = bs;
    pr->ginv = gi;
    pr->mode = md;
    pr->seeds = seeds;
    pr->W = W;
    pr->stats = stats;
    pr->flips = flips;
    pr->R = R;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R;
    pr->nthreads = nthreads;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    po_chain_job* jobs = (po_chain_job*)malloc(sizeof(po_chain_job) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].ch = &ch;
        jobs[t].t = t;
        pthread_create(&th[t], NULL, po_chain_worker, &jobs[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th); free(jobs); free(bs); free(gi); free(md);
    free(pr->rowptr); free(pr->col); free(pr->J); free(ch.chain_off); free(ch.chain_list);
    return 0;
}

/* The f32 helpers, exposed so the tests can check them against libm. */
float orc_pimc_exp2f(float y) { return po_exp2f(y); }
float orc_pimc_log2u(uint32_t v) { return po_log2u(v); }
uint32_t orc_pimc_xoshiro(uint32_t* s, int count, uint32_t* out) {
    po_rng g = {s[0], s[1], s[2], s[3]};
    for (int k = 0; k < count; ++k) out[k] = po_next(&g);
    s[0] = g.s0;
    s[1] = g.s1;
    s[2] = g.s2;
    s[3] = g.s3;
    return 0;
}
