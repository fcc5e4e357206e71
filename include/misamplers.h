/*
 * misamplers — MI355X (gfx950) Monte Carlo sweep engine, C ABI.
 *
 * Plain C types only (pointers and sizes); no exceptions cross this boundary.
 * Every function returns MISA_OK (0) or a negative MISA_ERR_* code, and
 * misa_last_error() gives the message of the calling thread's last failure.
 * Host-pointer entry points leave all buffers owned by the caller; device
 * buffers made by the library are owned by the problem / plan objects.
 *
 * Threads: the library keeps no process-global mutable state (the error
 * message is per thread, device attributes are queried per call).  A problem
 * may be shared by any number of threads and plans: it is read-only after
 * creation, and its device copies are made once, under the problem's lock.
 * A plan, batch or moments object is used by one thread at a time: it holds
 * per-plan buffers, captured graphs and the last call's statistics, so calls
 * on one plan that overlap race.  Calls on different plans are independent;
 * device-buffer calls run on the stream the caller passes.
 *
 * Reference interface replaced (paths relative to the dwave-samplers tree):
 *   general_simulated_annealing   dwave/samplers/sa/src/cpu_sa.h:63-78
 *                                 (definition cpu_sa.cpp:253-362), bound by
 *                                 dwave/samplers/sa/simulated_annealing.pyx:28-42
 *   general_simulated_annealing   dwave/samplers/sqa/src/localPIMC.h:60-86
 *   (path-integral)               (definition localPIMC.cpp:604-715), bound by
 *                                 dwave/samplers/sqa/pimc_annealing.pyx:30-55
 *   steepest_gradient_descent     dwave/samplers/greedy/src/descent.h (definition
 *                                 descent.cpp:337-406), bound by
 *                                 dwave/samplers/greedy/descent.pyx:97-101
 *   general_simulated_annealing   dwave/samplers/sqa/src/cpu_rotormc.h:63-82
 *   (rotor model)                 (definition cpu_rotormc.cpp:366-562), bound by
 *                                 dwave/samplers/sqa/rotormc_annealing.pyx:27-50
 *   sampleBQM (tree decomposition) dwave/samplers/tree/src/include/sample.hpp:244-276,
 *                                 bound by dwave/samplers/tree/sample.pyx:32-48
 */
#ifndef MISAMPLERS_H
#define MISAMPLERS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MISA_OK 0
#define MISA_ERR_ARG (-1)          /* bad argument (sizes, enums, nulls) */
#define MISA_ERR_LENGTHS (-2)      /* "coupler vectors have mismatched lengths" (cpu_sa.cpp:276-279) */
#define MISA_ERR_INDEX (-3)        /* "coupler indexes contain an invalid variable" (cpu_sa.cpp:302-304) */
#define MISA_ERR_DEVICE (-4)       /* HIP runtime error or no usable gfx950 device */
#define MISA_ERR_UNSUPPORTED (-5)  /* problem outside the engine's limits */

/* Enum values equal the reference's (cpu_sa.h:36-43). */
#define MISA_GIBBS 0
#define MISA_METROPOLIS 1
#define MISA_SEQUENTIAL 0
#define MISA_RANDOM 1

/* Value modes chosen per problem (MISA_MODE_AUTO picks the narrowest exact one). */
#define MISA_MODE_AUTO 0
#define MISA_MODE_INT8 1   /* integer biases, |field| <= 63: packed int8 state, host-built acceptance table */
#define MISA_MODE_INT16 2  /* integer biases, |field| <= 255: packed int16 state, host-built table */
#define MISA_MODE_REAL 3   /* float64 fields, device exp */

typedef struct misa_problem misa_problem;
typedef struct misa_sa_plan misa_sa_plan;

/* Called between completed reads, in read order; non-zero stops sampling
 * (sa/simulated_annealing.pyx:192-197 semantics: the call returns the number
 * of reads completed up to and including the one after which it returned
 * non-zero). */
typedef int (*misa_interrupt_cb)(void* ctx);

typedef struct misa_sa_stats {
    uint64_t flips;       /* accepted spin flips over all completed reads */
    uint64_t flip_degree; /* sum of deg(v) over those flips (neighbour updates) */
    double kernel_ms;     /* device time of the sweep launches */
    int mode;             /* MISA_MODE_* actually used */
    int lanes_per_read;   /* lanes cooperating on one read */
    int reads_per_block;
    int state_in_lds;     /* 1: per-read state in LDS, 0: in device memory */
    int kernel;           /* 1: integer LDS kernel (sa_int_kernel), 2: float64 LDS kernel
                             (sa_real_kernel), 3: batched float64 kernel (sa_batch_kernel),
                             4: float64 run kernel (sa_seg_kernel: non-adjacent consecutive
                             variables decided together), 5: int8 chain-run kernel
                             (sa_chain_kernel: runs of consecutive variables coupled only to
                             their predecessor decided together), 6: float64 dense kernel
                             (sa_dense_kernel: n <= 128, fields in registers, couplings as a
                             dense matrix in LDS), 0: general kernel (sa_sweep_kernel) */
    int csr_in_lds;       /* 1: adjacency staged in LDS */
    int waves_per_block;
    int chain_form;       /* kernel 5 only: 1 generator stepped in each run, 2 per-read output
                             rings with per-lane comparisons, 3 output rings with one ballot per
                             distinct threshold (few-threshold form); 0 otherwise */
} misa_sa_stats;

const char* misa_last_error(void);
const char* misa_version(void);
/* Number of visible devices (0 when none; never an error). */
int misa_device_count(void);

/* Problem: an Ising model h (num_vars) plus couplers (starts, ends, weights)
 * in the reference's raw form; self-loops and repeated couplers are allowed
 * (tests/test_simulated_annealing.py:98-100).  Adjacency keeps coupler
 * insertion order (cpu_sa.cpp:298-316).  Device copies are made lazily. */
int misa_problem_create(int num_vars, const double* h,
                        int64_t num_starts, const int32_t* coupler_starts,
                        int64_t num_ends, const int32_t* coupler_ends,
                        int64_t num_weights, const double* coupler_weights,
                        misa_problem** out);
void misa_problem_destroy(misa_problem* p);
int misa_problem_num_vars(const misa_problem* p);
int64_t misa_problem_num_couplers(const misa_problem* p);
int misa_problem_max_degree(const misa_problem* p);
/* The mode MISA_MODE_AUTO resolves to for this problem. */
int misa_problem_auto_mode(const misa_problem* p);

/* Host-buffer sampling (replaces general_simulated_annealing, cpu_sa.h:63-78).
 * states: (num_samples x num_vars) int8 spins, in = initial, out = samples.
 * energies: num_samples doubles (out).  read_seeds: one xorshift128+ seed per
 * read; read r behaves as the reference called with num_samples=1 and
 * seed=read_seeds[r].  Returns the number of completed reads (>= 0), or a
 * negative MISA_ERR_* code.  mode/lanes_per_read: 0 = automatic.  stats may
 * be NULL; the flip counts are only kept when it is not. */
int misa_sa_sample(const misa_problem* p, int8_t* states, double* energies,
                   int num_samples, int sweeps_per_beta,
                   const double* beta_schedule, int num_betas,
                   const uint64_t* read_seeds, int varorder, int proposal,
                   int mode, int lanes_per_read,
                   misa_interrupt_cb interrupt, void* interrupt_ctx,
                   misa_sa_stats* stats);

/* Plan: schedule-dependent tables uploaded once, for repeated device runs. */
int misa_sa_plan_create(const misa_problem* p, int sweeps_per_beta,
                        const double* beta_schedule, int num_betas, int varorder,
                        int proposal, int mode, int lanes_per_read,
                        misa_sa_plan** out);
void misa_sa_plan_destroy(misa_sa_plan* plan);
int misa_sa_plan_info(const misa_sa_plan* plan, misa_sa_stats* info);

/* Device-buffer run on a HIP stream (hipStream_t passed as void*, NULL = the
 * null stream).  d_states (num_samples x num_vars int8), d_energies
 * (num_samples f64) and d_seeds (num_samples u64) are device pointers;
 * d_counters, if not NULL, receives per-read {flips, flip_degree} u64 pairs
 * (NULL skips the counting; the samples are the same either way).
 * Asynchronous: returns after enqueueing. */
int misa_sa_run_device(misa_sa_plan* plan, int8_t* d_states, double* d_energies,
                       const uint64_t* d_seeds, int num_samples,
                       uint64_t* d_counters, void* stream);

/* misa_sa_run_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during a
 * run (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as
 * misa_sa_run_device; the one buffer that call may allocate, the per-read
 * state of reads that do not fit LDS (general kernel), lives instead in a
 * library-owned static device array of 64 MiB per GPU, held until the
 * process ends.  Runs needing more return MISA_ERR_UNSUPPORTED.  Every call on
 * a GPU shares that GPU's array: calls that overlap in time on one GPU race.
 * The plan (created beforehand) holds every other table the run uses. */
int misa_sa_run_device_noalloc(misa_sa_plan* plan, int8_t* d_states, double* d_energies,
                               const uint64_t* d_seeds, int num_samples, uint64_t* d_counters,
                               void* stream);

/* ---- Batched launch (many problems of uneven size in one launch) --------
 * No reference counterpart: the reference samples one BQM per call.  All
 * problems share the variable order and proposal; each has its own reads and
 * schedule.  The problems must outlive the batch.  States / energies / seeds
 * are concatenated in problem order:
 * problem b's reads occupy states[state_offset(b) ...] (R_b x n_b int8) and
 * energies/seeds[read_offset(b) ...] (R_b).  The batch runs the float64
 * kernel (integer problems included). */
typedef struct misa_batch misa_batch;

int misa_batch_create(int num_problems, const misa_problem* const* problems, const int* num_reads,
                      const int* sweeps_per_beta, const double* const* beta_schedules,
                      const int* num_betas, int varorder, int proposal, misa_batch** out);
void misa_batch_destroy(misa_batch* batch);
int64_t misa_batch_state_offset(const misa_batch* batch, int problem);
int64_t misa_batch_read_offset(const misa_batch* batch, int problem);
int64_t misa_batch_total_reads(const misa_batch* batch);
int64_t misa_batch_total_state(const misa_batch* batch);
/* Geometry policy, 0..15: bit 0 stages the adjacency in LDS first (else only
 * when it costs few reads per block), bit 1 aims at more waves per block
 * (float64 row body: half the reads per block), bits 2-3 cap the reads per
 * wave of the run body at 8 >> bits (8, 4, 2 or 1: at least 8, 16, 32 or 64
 * lanes per read).  misa_batch_autotune times every policy on `trial_sweeps`
 * sweeps (device buffers of its own) and keeps the fastest; returns it. */
int misa_batch_set_policy(misa_batch* batch, int policy);
int misa_batch_autotune(misa_batch* batch, int trial_sweeps);
/* Device-buffer run on a HIP stream.  With use_graph != 0 the launch is
 * captured once per buffer set into a hipGraph and replayed (stream must not
 * be the null stream). */
int misa_batch_run_device(misa_batch* batch, int8_t* d_states, double* d_energies, const uint64_t* d_seeds,
                          uint64_t* d_counters, void* stream, int use_graph);
/* Host-buffer convenience: copies in, runs, copies out.  Returns MISA_OK. */
int misa_batch_sample(misa_batch* batch, int8_t* states, double* energies, const uint64_t* seeds,
                      misa_sa_stats* stats);
/* New biases for every problem of the batch, read from device memory on
 * `stream`: d_h (sum of num_vars float32, problems in order) and d_J (sum of
 * the coupler counts float32, each problem's couplers in input order), widened
 * to float64.  No reference counterpart: it serves callers whose model values
 * live on the GPU (differentiable fixed-beta sampling), with no host copy or
 * host synchronisation; the coupler structure stays as created.  The first
 * call allocates the index maps; later calls allocate nothing.  The batch's
 * problems then hold these values on the device only: use them through this
 * batch alone.  A batch launch still running on another stream is waited for
 * on the host. */
int misa_batch_set_values_device(misa_batch* batch, const float* d_h, const float* d_J, void* stream);

/* ---- Path-integral (transverse-field) annealing ---------------------------
 * Discrete imaginary time: every read carries P <= 32 Trotter slices per
 * variable, one bit each of a uint32 worldline word (bit k = slice k, set =
 * spin +1).  The sampled law is the Trotter model
 *   Z_P = Tr[(exp(-(Hp/P) Hz) exp((Hd/P) sum_i X_i))^P],
 *   Hz  = sum_i h_i Z_i + sum_c J_c Z_u Z_v,
 * which tends to the reference's continuous-time model (localPIMC.cpp:17-21,
 * 2^16 slices) as P grows.  Self-loops are constants (s^2 = 1).  Sites are
 * visited colour class by colour class; each (read, sweep, site) update uses
 * its own random stream, so results depend on the seeds only. */
typedef struct misa_pimc_plan misa_pimc_plan;

typedef struct misa_pimc_stats {
    double kernel_ms;       /* device time of the sweep launches */
    int64_t site_updates;   /* worldline (site) updates: reads x sweeps x vars; with
                               chain updates reads x sweeps x Q x (chains x Q) */
    int num_slices;         /* P */
    int reads_per_block;
    int threads_per_block;
    int num_classes;        /* colour classes per sweep */
    int kind;               /* 1: all |J| equal (sign-folded popcounts), 0: coupling table,
                               3: chain updates (coupling table) */
    int lds_bytes;
    int max_degree;         /* largest row of the adjacency without self-loops */
    int blocks;
} misa_pimc_stats;

/* Plan: adjacency without self-loops, colour classes and per-step tables
 * (Hp/P and the cut law of tanh(Hd/P)) for hp_field/hd_field[0..num_fields).
 * num_slices in [1, 32].  MISA_ERR_UNSUPPORTED beyond the engine's limits
 * (more than 65535 variables, or 32767 when all |J| are equal). */
int misa_pimc_plan_create(const misa_problem* p, int sweeps_per_beta, const double* hp_field,
                          const double* hd_field, int num_fields, int num_slices, misa_pimc_plan** out);
/* Plan of chain updates (replaces the reference's chainUpdate path,
 * localPIMC.cpp:95-126 with qubitsPerUpdate != 1, :146-152, :385-558): chain
 * c holds qubits c*Q .. c*Q+Q-1 for c < num_vars / Q (later qubits are never
 * updated), consecutive qubits of a chain coupled by the implicit
 * chain_coupler_strength (not in the coupler list).  One chain update cuts
 * every qubit's worldline as a site update does, joins the segments of
 * consecutive qubits piece by piece where their spins agree with probability
 * 1 - exp(2 Jchain (Hp/P) len) (Jchain >= 0 never joins), and draws one
 * heat-bath spin per component.  A sweep makes Q passes over the chains,
 * colour class by colour class of the chain graph.  qubits_per_chain in
 * [1, 32] with qubits_per_chain x num_slices <= 512 (MISA_ERR_UNSUPPORTED
 * beyond). */
int misa_pimc_plan_create_chains(const misa_problem* p, int sweeps_per_beta, const double* hp_field,
                                 const double* hd_field, int num_fields, int num_slices, int qubits_per_chain,
                                 double chain_coupler_strength, misa_pimc_plan** out);
void misa_pimc_plan_destroy(misa_pimc_plan* plan);
/* The colour class of every variable (num_vars int32) in the engine's update
 * order: classes are visited in increasing colour, a class's variables are
 * pairwise non-adjacent.  Host-only (no device needed).  Returns the number
 * of classes. */
int misa_pimc_colouring(const misa_problem* p, int32_t* colour_out);
/* The colour class of every chain (num_vars / qubits_per_chain int32) in the
 * chain-update order of misa_pimc_plan_create_chains: chains of one class
 * share no coupler.  Host-only.  Returns the number of classes. */
int misa_pimc_chain_colouring(const misa_problem* p, int qubits_per_chain, int32_t* colour_out);
/* Launch shape the plan would use for num_reads reads. */
int misa_pimc_plan_info(const misa_pimc_plan* plan, int num_reads, misa_pimc_stats* info);

/* Device-buffer run: d_worldlines (num_reads x num_vars uint32) in/out,
 * d_seeds (num_reads u64), d_statistics (num_reads x (1 + (num_fields-1) /
 * interval) x num_vars int8) receives slice 0 after every interval-th step
 * when interval > 0 (may be NULL when interval == 0); d_flips, if not NULL,
 * (num_reads u64, caller-zeroed) accumulates the number of slice spins each
 * read changed.  Asynchronous. */
int misa_pimc_run_device(misa_pimc_plan* plan, uint32_t* d_worldlines, const uint64_t* d_seeds,
                         int num_reads, int8_t* d_statistics, int schedule_sample_interval,
                         uint64_t* d_flips, void* stream);
/* Slice 0 of each read as +-1 spins and its energy in the reference's order
 * (localPIMC.cpp:56-73); either output may be NULL. */
int misa_pimc_project_device(const misa_problem* p, const uint32_t* d_worldlines, int num_reads,
                             int8_t* d_states, double* d_energies, void* stream);

/* Host-buffer sampling (replaces the path-integral general_simulated_annealing,
 * localPIMC.h:60-86, argument for argument, plus read_seeds / num_slices):
 *   states (num_samples x num_vars int8) in/out: initial classical states or
 *     first slices, then the projected samples (slice 0);
 *   project_inputs != 0: worldlines start constant in imaginary time;
 *     otherwise num_breaks (num_samples x num_vars) and the flat breaks_in
 *     (breaks_in_size entries, positions in [0, 2^16)) give the initial
 *     worldlines as in localPIMC.cpp:169-180;
 *   project_outputs == 0: num_breaks and breaks_buffer_out
 *     (breaks_buffer_size entries) receive the final worldlines in the
 *     reference's format and buffer discipline (localPIMC.cpp:188-199);
 *     a kink between slices k-1 and k is a break at (2k-1) * 2^16 / (2P);
 *   gamma is unused, as in the reference (run() overwrites it with
 *     hd_field, localPIMC.cpp:136); qubits_per_update == 1 selects
 *     site updates (chain_coupler_strength unused, as in the reference's
 *     generic constructor), any other value chain updates of
 *     qubits_per_chain qubits (see misa_pimc_plan_create_chains);
 *     qubits_per_chain >= qubits_per_update (localPIMC.cpp:101);
 *   statistics: see misa_pimc_run_device (may be NULL when interval == 0).
 * Returns the number of completed reads, or a negative MISA_ERR_* code. */
int misa_pimc_sample(const misa_problem* p, int8_t* states, double* energies, int num_samples,
                     int project_inputs, int project_outputs, int32_t* num_breaks, const int32_t* breaks_in,
                     int64_t breaks_in_size, int32_t* breaks_buffer_out, int64_t breaks_buffer_size,
                     int sweeps_per_beta, const double* hp_field, const double* hd_field, int num_fields,
                     double gamma, double chain_coupler_strength, int qubits_per_chain,
                     int qubits_per_update, const uint64_t* read_seeds, int num_slices,
                     int8_t* statistics, int schedule_sample_interval, misa_interrupt_cb interrupt,
                     void* interrupt_ctx, misa_pimc_stats* stats);

/* ---- Steepest descent (greedy local search) -------------------------------
 * Replaces steepest_gradient_descent (descent.cpp:337-406): every read
 * descends from its initial state, flipping the variable with the most
 * negative flip energy (lowest index on ties) until none is negative, with
 * the reference's float64 bookkeeping (descent.cpp:117-187).  Self-loops are
 * constant energy (the reference's bookkeeping for them can cycle forever).
 * A read stops after max_steps flips (0xFFFFFFFF: no practical limit). */
int misa_descent_run_device(const misa_problem* p, int8_t* d_states, double* d_energies,
                            uint32_t* d_steps, int num_samples, uint32_t max_steps, void* stream);
/* Host buffers: states (num_samples x num_vars int8) in/out, energies and
 * num_steps out; kernel_ms (may be NULL) receives the device time. */
int misa_descent_sample(const misa_problem* p, int8_t* states, double* energies, uint32_t* num_steps,
                        int num_samples, uint32_t max_steps, double* kernel_ms);

/* ---- Rotor-model annealing ------------------------------------------------
 * Replaces general_simulated_annealing of the rotor model
 * (dwave/samplers/sqa/src/cpu_rotormc.cpp:366-562, declared in
 * cpu_rotormc.h:63-82).  States are angles i in [0, 256) (theta = 2 pi i /
 * 256), num_samples x num_vars uint8, in/out.  Read r is the reference core
 * called with num_samples = 1, seed = read_seeds[r] and read r's initial
 * angles.  proposal: the reference's enum, 0 GibbsNonErgodic,
 * 1 MetropolisNonErgodic, 2 MetropolisUniform, 3 MetropolisTF.  Energies are
 * the cos-weighted classical energies (cpu_rotormc.cpp:310-329).  statistics
 * (may be NULL, or interval 0): num_samples x (1 + (num_fields - 1) /
 * interval) x num_vars angles, a snapshot after every schedule step whose
 * index is a multiple of the interval (rotormc_annealing.pyx:171-176). */
#define MISA_ROTOR_GIBBS_NONERGODIC 0
#define MISA_ROTOR_METROPOLIS_NONERGODIC 1
#define MISA_ROTOR_METROPOLIS_UNIFORM 2
#define MISA_ROTOR_METROPOLIS_TF 3

typedef struct misa_rotor_plan misa_rotor_plan;

typedef struct misa_rotor_stats {
    double kernel_ms;       /* device time of the sweep launches */
    int64_t accepted;       /* accepted proposals over all completed reads */
    int lanes_per_read;
    int reads_per_block;
    int threads_per_block;
    int csr_in_lds;         /* adjacency staged in LDS */
    int lds_bytes;
    int wave_per_read;      /* sequential angle proposals: one wave per read */
    int run_kernel;         /* uniform proposals decided in runs of pairwise non-adjacent variables */
} misa_rotor_stats;

/* Plan: the transverse fields (num_vars), the schedule (Hp = inverse
 * temperature, Hd = transverse-field scale, num_fields steps of
 * sweeps_per_beta sweeps), the order and the proposal, uploaded once. */
int misa_rotor_plan_create(const misa_problem* p, const double* trans_fields, int sweeps_per_beta,
                           const double* hp_field, const double* hd_field, int num_fields, int randomize_order,
                           int proposal, misa_rotor_plan** out);
void misa_rotor_plan_destroy(misa_rotor_plan* plan);
int misa_rotor_plan_info(const misa_rotor_plan* plan, misa_rotor_stats* info);
/* Device buffers: d_states (num_samples x num_vars angles) in/out,
 * d_energies, d_seeds, d_statistics (may be NULL), d_accepted (per read, may
 * be NULL); enqueued on `stream`. */
int misa_rotor_run_device(const misa_rotor_plan* plan, uint8_t* d_states, double* d_energies,
                          const uint64_t* d_seeds, int num_samples, uint8_t* d_statistics,
                          int schedule_sample_interval, unsigned long long* d_accepted, void* stream);
/* Host buffers; the interrupt callback is polled after every completed read,
 * in read order.  Returns the number of completed reads, or a negative
 * MISA_ERR_* code. */
int misa_rotor_sample(const misa_problem* p, uint8_t* states, double* energies, int num_samples,
                      const double* trans_fields, int sweeps_per_beta, const double* hp_field,
                      const double* hd_field, int num_fields, const uint64_t* read_seeds, int randomize_order,
                      int proposal, uint8_t* statistics, int schedule_sample_interval, misa_interrupt_cb interrupt,
                      void* interrupt_ctx, misa_rotor_stats* stats);

/* ---- Exact inference by bucket elimination (tree decomposition) ----------
 * Counterpart of the reference's TreeDecompositionSampler core
 * (dwave/samplers/tree/samplers.py:235-359, tree/sample.pyx:48-175): the
 * Boltzmann distribution P(x) ~ exp(-beta E(x)) of the problem, x_i in
 * {low, 1} (low = -1 spin, 0 binary).  Repeated couplers are added; a
 * self-loop is a constant (spin) or a field (binary).  The plan fixes the
 * elimination order (NULL: min-fill) and allocates the bucket tables
 * (2^(width+1) doubles per bucket at most); widths above max_treewidth give
 * MISA_ERR_UNSUPPORTED. */
typedef struct misa_tree_plan misa_tree_plan;

int misa_tree_plan_create(const misa_problem* p, const int32_t* elimination_order, double beta, int low,
                          int max_treewidth, misa_tree_plan** out);
void misa_tree_plan_destroy(misa_tree_plan* plan);
/* Width of the order, table entries, number of merged couplers (pairs), and
 * the order itself (num_vars ints; any argument may be NULL). */
int misa_tree_plan_info(const misa_tree_plan* plan, int* treewidth, int64_t* table_entries, int* num_pairs,
                        int32_t* elimination_order);
/* The next marginals or samples call evaluates the tables again (they are
 * otherwise computed once per plan). */
int misa_tree_recompute(misa_tree_plan* plan);
/* The merged couplers (u, v) in first-appearance order. */
int misa_tree_pairs(const misa_tree_plan* plan, int32_t* u, int32_t* v);
/* log Z (log partition function, without any energy offset), P(x_i = 1) per
 * variable (num_vars, may be NULL) and per pair the four probabilities
 * P(x_u, x_v) in the order (low, low), (1, low), (low, 1), (1, 1) (4 per pair,
 * may be NULL). */
int misa_tree_marginals(misa_tree_plan* plan, double* log_z, double* var_p1, double* pair_p, void* stream);
/* The k lowest-energy states (replaces TreeDecompositionSolver's
 * solve_bqm_wrapper, dwave/samplers/tree/samplers.py:166-170, and its k-best
 * bucket elimination in dwave/samplers/tree/src): min-sum elimination in the
 * plan's order with lists of the k smallest energies per separator
 * assignment, then backtracking.  k = min(max_solutions, 2^num_vars).
 * samples (k x num_vars int8; 1 / -1 for spin plans, 1 / 0 for binary) and
 * energies (ascending, the problem's energy without its offset; ties in a
 * fixed order) receive *num_found states.  The plan's beta is not used. */
int misa_tree_solve(misa_tree_plan* plan, int max_solutions, int8_t* samples, double* energies, int* num_found);

/* num_samples exact samples (num_samples x num_vars int8 of low / 1 values),
 * read r from its own xorshift128+ stream whose two state words are
 * splitmix64 mixes of read_seeds[r]. */
int misa_tree_sample(misa_tree_plan* plan, int num_samples, const uint64_t* read_seeds, int8_t* samples,
                     void* stream);
/* Device time of the last call's passes (HIP events on the caller's stream):
 * the upward pass and log Z (0 when the tables were current), the downward
 * pass and marginals (0 when none were asked), and the work they ran. */
typedef struct misa_tree_stats {
    double forward_ms;
    double marginals_ms;
    int forward_levels;      /* launches of the upward pass */
    int downward_levels;     /* launches of the downward pass */
    int64_t marginal_blocks; /* chunks of the marginal reduction */
    int64_t table_entries;   /* entries of all T tables */
    double forward_bytes;    /* algorithmic bytes of each pass: table entries */
    double downward_bytes;   /* written and read once per use */
    double marginal_bytes;
} misa_tree_stats;
int misa_tree_last_stats(const misa_tree_plan* plan, misa_tree_stats* out);

/* ---- Spin moments (backward pass of fixed-beta sampling) -----------------
 * No reference counterpart: it serves differentiable sampling, where
 * F = -(1/beta) log Z with E = sum_i h_i s_i + sum_c J_c s_u s_v has
 * dF/dh_i = <s_i> and dF/dJ_c = <s_u s_v>.  Models b = 0..num_models-1 are
 * concatenated in order: samples (num_reads[b] x num_vars[b] int8 +-1, row
 * major), edges ((u, v) int32 pairs, num_edges[b] of them), grad_h
 * (num_vars[b] floats), grad_J (num_edges[b] floats), sums (num_vars[b] +
 * num_edges[b] int64: exact sum_r s_i, then sum_r s_u s_v).  Every edge must
 * index variables of its own model (checked by the caller: the kernel does not
 * bounds-check them).  The plan owns its
 * partial-sum buffer: runs allocate nothing, and two runs of one plan must not
 * overlap in time. */
typedef struct misa_moments_plan misa_moments_plan;

int misa_moments_plan_create(int num_models, const int* num_vars, const int64_t* num_edges,
                             const int* num_reads, misa_moments_plan** out);
void misa_moments_plan_destroy(misa_moments_plan* plan);
/* grad_h[i] = scale[b] * sum / num_reads[b] (float32), likewise grad_J;
 * d_scale may be NULL (1), any output may be NULL.  Asynchronous. */
int misa_moments_run(misa_moments_plan* plan, const int8_t* d_samples, const int32_t* d_edges,
                     const float* d_scale, float* d_grad_h, float* d_grad_J, int64_t* d_sums,
                     void* stream);

/* Measured LDS read bandwidth of the current device in GB/s (16-byte
 * conflict-free reads on every CU, `iters` per thread; about 4096 gives a
 * launch of a few milliseconds), and the kernel time.  No reference
 * counterpart: the peak that LDS-resident kernels are priced against. */
int misa_lds_bandwidth(int iters, double* gbs, double* ms);

/* The engine's per-read seed derivation: out[r] = splitmix64 output r of `seed`. */
void misa_read_seeds(uint64_t seed, int64_t num_reads, uint64_t* out);
/* misa_read_seeds into device memory, asynchronously on `stream`. */
int misa_read_seeds_device(uint64_t seed, int64_t num_reads, uint64_t* d_out, void* stream);
/* Random +-1 spins: out[i] = +1 when the top bit of splitmix64 output i of
 * (seed XOR 0xD1B54A32D192ED03) is set, else -1.  No reference counterpart:
 * the initial states of differentiable fixed-beta sampling, drawn on the
 * device without a host copy (the _device form) and reproducible on the host. */
void misa_random_spins(uint64_t seed, int64_t count, int8_t* out);
int misa_random_spins_device(uint64_t seed, int64_t count, int8_t* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MISAMPLERS_H */
