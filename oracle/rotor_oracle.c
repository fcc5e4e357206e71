/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this; the engine never links it.
 *
 * Plain-C restatement of the reference's rotor-model annealing core
 * (dwave/samplers/sqa/src/cpu_rotormc.cpp), one read per call of the
 * reference with num_samples = 1 and that read's seed:
 *   FASTRAND xorshift128+, seeding {seed ? seed : ~0, 0}  cpu_rotormc.cpp:31-38, 407-408
 *   angle tables cos(2 i pi / 256), |sin(2 i pi / 256)|    :392-396
 *   get_marginal_state_fieldC (sum in adjacency order)     :58-79
 *   simulated_annealing_run<order, proposal>               :106-297
 *       32 burn-in draws (:145-147); per schedule step beta = Hp, beta_tf = Hd,
 *       rat = beta_tf / beta (:153-157); per variable (sequential, or
 *       rand % n) one of
 *         MetropolisNonErgodic  reflect theta -> theta + 128 (:172-190)
 *         GibbsNonErgodic       reflect, Gibbs acceptance     (:191-202)
 *         MetropolisUniform     uniform angle rand % 256      (:203-206)
 *         MetropolisTF          uniform if rat >= 1, else a window of
 *                               half-width osi(+1) around theta (:207-242)
 *       the last two accept on the log-measure change
 *       -beta f (cos' - cos) + beta_tf t (|sin'| - |sin|) (:246-266);
 *       a reflection updates the neighbours' dE (:268-285);
 *       snapshots after steps with index % interval == 0 (:290-294)
 *   get_state_energyC (cos-weighted, variable order then couplers) :310-329
 * The two integer steps of MetropolisTF that C++ leaves to the machine are
 * written out as the x86-64 build of the reference computes them: double ->
 * int truncation gives INT_MIN for NaN or out-of-range values (cvttsd2si),
 * and int arithmetic wraps.  Snapshots are laid out (read, 1 + (len-1) /
 * interval, n), the shape the Cython layer allocates (rotormc_annealing.pyx:
 * 171-176).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { RO_GIBBS_NE = 0, RO_METROPOLIS_NE = 1, RO_UNIFORM = 2, RO_TF = 3 };

static const double kTwo64 = 18446744073709551616.0; /* (double)RANDMAX */

static inline uint64_t ro_next(uint64_t st[2]) {
    uint64_t x = st[0];
    const uint64_t y = st[1];
    st[0] = y;
    x ^= x << 23;
    st[1] = x ^ y ^ (x >> 17) ^ (y >> 26);
    return st[1] + y;
}

/* cvttsd2si: truncation toward zero, INT32_MIN for NaN and out-of-range. */
static inline int32_t ro_trunc_i32(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}

void orc_rotor_tables(double* cos_t, double* sin_t) {
    /* cpu_rotormc.cpp:392-395: cos(2*index*M_PI/discretization) */
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < 256; ++i) {
        cos_t[i] = cos(2 * i * pi / 256);
        sin_t[i] = fabs(sin(2 * i * pi / 256));
    }
}

typedef struct {
    int n;
    const double* h;
    const int* rowptr;
    const int* col;
    const double* J;
    int64_t m;
    const int* cs;
    const int* ce;
    const double* cw;
    const double* tf;
    int spb;
    const double* hp;
    const double* hd;
    int nf;
    const uint64_t* seeds;
    int randomize;
    int proposal;
    uint8_t* states;
    double* energies;
    uint8_t* stats;
    int interval;
    int64_t stat_block;
    double cos_t[256], sin_t[256];
} RoJob;

static double ro_field(const RoJob* jb, int v, const uint8_t* st) {
    double e = jb->h[v];
    for (int k = jb->rowptr[v]; k < jb->rowptr[v + 1]; ++k) e += jb->cos_t[st[jb->col[k]]] * jb->J[k];
    return e;
}

static void ro_read(const RoJob* jb, int r, double* dE) {
    const int n = jb->n;
    uint8_t* st = jb->states + (size_t)r * n;
    uint8_t* stats = jb->stats ? jb->stats + (size_t)r * jb->stat_block : NULL;
    const double* ct = jb->cos_t;
    const double* sn = jb->sin_t;
    const int ne = jb->proposal == RO_GIBBS_NE || jb->proposal == RO_METROPOLIS_NE;
    uint64_t rs[2] = {jb->seeds[r] ? jb->seeds[r] : ~0ull, 0};
    for (int v = 0; v < n; ++v) dE[v] = ro_field(jb, v, st);
    if (ne)
        for (int v = 0; v < n; ++v) dE[v] *= -2 * ct[st[v]];
    uint64_t rnd;
    for (int i = 0; i < 32; ++i) rnd = ro_next(rs);
    for (int bi = 0; bi < jb->nf; ++bi) {
        const double beta = jb->hp[bi], beta_tf = jb->hd[bi];
        const double rat = beta_tf / beta;
        for (int sw = 0; sw < jb->spb; ++sw) {
            for (int it = 0; it < n; ++it) {
                int v = it;
                if (jb->randomize) {
                    rnd = ro_next(rs);
                    v = (int)(rnd % (uint64_t)n);
                }
                uint8_t prop = 0;
                double prop_energy = 0.0;
                int flip = 0;
                switch (jb->proposal) {
                    case RO_METROPOLIS_NE:
                        prop = (uint8_t)(128 + st[v]);
                        prop_energy = -dE[v];
                        if (dE[v] <= 0.0) {
                            flip = 1;
                        } else {
                            rnd = ro_next(rs);
                            if (exp(-dE[v] * beta) * kTwo64 > (double)rnd) flip = 1;
                        }
                        break;
                    case RO_GIBBS_NE:
                        prop = (uint8_t)(128 + st[v]);
                        prop_energy = -dE[v];
                        rnd = ro_next(rs);
                        if (kTwo64 > (double)rnd * (1 + exp(dE[v] * beta))) flip = 1;
                        break;
                    case RO_UNIFORM:
                        rnd = ro_next(rs);
                        prop = (uint8_t)(rnd % 256);
                        break;
                    default: /* RO_TF */
                        rnd = ro_next(rs);
                        if (rat >= 1) {
                            prop = (uint8_t)(rnd % 256);
                        } else {
                            const int32_t osi = ro_trunc_i32(rat * 128 - 0.5);
                            const int32_t prod = (int32_t)(uint32_t)((int64_t)osi * (int64_t)osi +
                                                                     (int64_t)osi);  /* osi*(osi+1), wrapped */
                            const int32_t span = (int32_t)(2u * (uint32_t)osi + 1u);  /* 2*osi+1, wrapped */
                            const double distC = prod / (double)span;
                            const double p = (rat / 2 * 128 - distC) / ((int32_t)((uint32_t)osi + 1u) - distC);
                            const uint64_t randL = rnd;
                            rnd = ro_next(rs);
                            uint64_t pac;
                            if (p * kTwo64 > (double)randL)
                                pac = (uint64_t)(int64_t)(int32_t)((uint32_t)osi + 1u) * (2 * (rnd % 2) - 1);
                            else
                                pac = rnd % (uint64_t)(int64_t)span - (uint64_t)(int64_t)osi;
                            prop = (uint8_t)(st[v] + (uint8_t)pac);
                        }
                        break;
                }
                if (!ne) {
                    prop_energy = ro_field(jb, v, st);
                    const double dlm = -beta * prop_energy * (ct[prop] - ct[st[v]]) +
                                       beta_tf * jb->tf[v] * (sn[prop] - sn[st[v]]);
                    if (dlm >= 0.0) {
                        flip = 1;
                    } else {
                        rnd = ro_next(rs);
                        if (exp(dlm) * kTwo64 > (double)rnd) flip = 1;
                    }
                }
                if (flip) {
                    if (ne) {
                        const double mult = 2 * (ct[st[v]] - ct[prop]);
                        for (int k = jb->rowptr[v]; k < jb->rowptr[v + 1]; ++k) {
                            const int u = jb->col[k];
                            dE[u] += mult * jb->J[k] * ct[st[u]];
                        }
                        dE[v] = prop_energy;
                    }
                    st[v] = prop;
                }
            }
        }
        if (stats && jb->interval != 0 && bi % jb->interval == 0) {
            memcpy(stats, st, (size_t)n);
            stats += n;
        }
    }
    double en = 0.0;
    for (int v = 0; v < n; ++v) en += ct[st[v]] * jb->h[v];
    for (int64_t c = 0; c < jb->m; ++c) en += ct[st[jb->cs[c]]] * jb->cw[c] * ct[st[jb->ce[c]]];
    jb->energies[r] = en;
}

typedef struct {
    const RoJob* jb;
    int r0, r1;
} RoSlice;

static void* ro_worker(void* arg) {
    const RoSlice* s = (const RoSlice*)arg;
    double* dE = (double*)malloc(sizeof(double) * (size_t)(s->jb->n > 0 ? s->jb->n : 1));
    for (int r = s->r0; r < s->r1; ++r) ro_read(s->jb, r, dE);
    free(dE);
    return NULL;
}

/* Returns 0, or -2 for a coupler index out of range (cpu_rotormc.cpp:426-428).
 * Adjacency: coupler c appends v to row u, then u to row v (:420-441). */
int orc_rotor_sample(int n, const double* h, int64_t m, const int* cs, const int* ce, const double* cw,
                     const double* tf, int R, int spb, const double* hp, const double* hd, int nf,
                     const uint64_t* seeds, int randomize, int proposal, uint8_t* states, double* energies,
                     uint8_t* stats, int interval, int nthreads) {
    for (int64_t c = 0; c < m; ++c)
        if (cs[c] < 0 || ce[c] < 0 || cs[c] >= n || ce[c] >= n) return -2;
    RoJob* jb = (RoJob*)calloc(1, sizeof(RoJob));
    int* rowptr = (int*)calloc((size_t)n + 1, sizeof(int));
    int* col = (int*)malloc(sizeof(int) * (size_t)(2 * m + 1));
    double* J = (double*)malloc(sizeof(double) * (size_t)(2 * m + 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    for (int64_t c = 0; c < m; ++c) {
        rowptr[cs[c] + 1]++;
        rowptr[ce[c] + 1]++;
    }
    for (int v = 0; v < n; ++v) rowptr[v + 1] += rowptr[v];
    memcpy(fill, rowptr, sizeof(int) * (size_t)(n + 1));
    for (int64_t c = 0; c < m; ++c) {
        col[fill[cs[c]]] = ce[c];
        J[fill[cs[c]]++] = cw[c];
        col[fill[ce[c]]] = cs[c];
        J[fill[ce[c]]++] = cw[c];
    }
    jb->n = n;
    jb->h = h;
    jb->rowptr = rowptr;
    jb->col = col;
    jb->J = J;
    jb->m = m;
    jb->cs = cs;
    jb->ce = ce;
    jb->cw = cw;
    jb->tf = tf;
    jb->spb = spb;
    jb->hp = hp;
    jb->hd = hd;
    jb->nf = nf;
    jb->seeds = seeds;
    jb->randomize = randomize;
    jb->proposal = proposal;
    jb->states = states;
    jb->energies = energies;
    jb->stats = stats;
    jb->interval = interval;
    jb->stat_block = (interval > 0 && nf > 0) ? (int64_t)n * (1 + (nf - 1) / interval) : 0;
    orc_rotor_tables(jb->cos_t, jb->sin_t);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R > 0 ? R : 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    RoSlice* sl = (RoSlice*)malloc(sizeof(RoSlice) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        sl[t].jb = jb;
        sl[t].r0 = (int)((int64_t)R * t / nthreads);
        sl[t].r1 = (int)((int64_t)R * (t + 1) / nthreads);
        pthread_create(&th[t], NULL, ro_worker, &sl[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(sl);
    free(fill);
    free(J);
    free(col);
    free(rowptr);
    free(jb);
    return 0;
}
