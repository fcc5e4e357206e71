// C-ABI host side of the rotor-model annealing engine (include/misamplers.h).
//
// Replaces the reference's general_simulated_annealing of the rotor model
// (dwave/samplers/sqa/src/cpu_rotormc.cpp:366-562, declared in cpu_rotormc.h:
// 63-82): the coupler validation and adjacency are the problem object's (same
// insertion order and messages), the angle tables are computed here exactly as
// the reference computes them (:392-396), and each read runs on the device
// with its own xorshift128+ stream seeded from read_seeds[r].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host.h"
#include "misamplers.h"
#include "rotor_engine.h"
#include "sa_engine.h"

using namespace misa_host;

struct misa_rotor_plan {
    misa_problem* p = nullptr;  // not owned
    int spb = 0, nf = 0, randomize = 0, proposal = 0;
    double* d_tf = nullptr;
    double* d_hp = nullptr;
    double* d_hd = nullptr;
    double* d_tables = nullptr;
    // geometry
    int lanes_log2 = 0, rpb = 1, rpb_log2 = 0, threads = 64, csr_in_lds = 0, wave_per_read = 0;
    unsigned off_rowinfo = 0, off_col = 0, off_J = 0, off_dE = 0, off_ang = 0, off_runs = 0, off_jump = 0,
             lds_bytes = 0;
    // rotor_run_kernel: runs of consecutive pairwise non-adjacent variables
    int run_kernel = 0;
    std::vector<uint32_t> runs;
    uint32_t* d_runs = nullptr;
    uint32_t* d_run_rows = nullptr;
    const uint32_t* d_jump = nullptr;  // per device, not owned
    ~misa_rotor_plan() {
        dev_free(d_tf);
        dev_free(d_hp);
        dev_free(d_hd);
        dev_free(d_tables);
        dev_free(d_runs);
        dev_free(d_run_rows);
    }
};

namespace {

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

int log2_exact(int x) {
    int l = 0;
    while ((1 << l) < x) ++l;
    return l;
}

// Runs of rotor_run_kernel: consecutive variables, at most kRotorRunMax, no
// two adjacent (a variable ends the run when its row names an earlier
// member; its own entry, a self-loop, does not), as first | length << 16.
std::vector<uint32_t> rotor_runs(const misa_problem* p) {
    std::vector<uint32_t> runs;
    const int n = p->n;
    for (int s = 0; s < n;) {
        int e = s + 1;
        while (e < n && e - s < misa::kRotorRunMax) {
            bool clash = false;
            for (int k = p->rowptr[e]; k < p->rowptr[e + 1] && !clash; ++k)
                clash = p->col[k] >= s && p->col[k] < e;
            if (clash) break;
            ++e;
        }
        runs.push_back((uint32_t)s | (uint32_t)(e - s) << 16);
        s = e;
    }
    return runs;
}

// Reads per block, lanes per read and LDS layout.  The adjacency is staged in
// LDS whenever it fits next to at least one read (every attempt of the angle
// proposals re-reads a whole row); the rest of LDS holds as many reads as fit,
// rounded down to a power of two.  Reflection proposals get up to
// min(16, max degree) lanes per read for the neighbour updates; angle
// proposals use one lane per read unless LDS holds fewer than 64 reads.
int rotor_geometry(misa_rotor_plan* pl) {
    const misa_problem* p = pl->p;
    const int n = p->n;
    const bool ne = pl->proposal == misa::ROTOR_GIBBS_NE || pl->proposal == misa::ROTOR_METROPOLIS_NE;
    const size_t lmax = misa::max_dynamic_lds();
    const size_t m2 = 2 * (size_t)p->m;
    auto layout = [&](int csr, int rpb) -> size_t {
        size_t off = 4096;
        pl->off_rowinfo = (unsigned)off;
        off += align16(4 * ((size_t)n + 1));
        pl->off_col = (unsigned)off;
        if (csr) off += align16(2 * m2);
        pl->off_J = (unsigned)off;
        if (csr) off += align16(8 * m2);
        pl->off_dE = (unsigned)off;
        if (ne) off += align16(8 * (size_t)n * rpb);
        pl->off_ang = (unsigned)off;
        off += align16((size_t)n * rpb);
        pl->off_runs = (unsigned)off;
        pl->off_jump = (unsigned)off;
        return off;
    };
    // the run kernel's layout: the wave kernel's, then the runs and the jump table
    // the run kernel's layout: tables, row info, the adjacency with 16 zero
    // entries after it (rows are read past their end), four reads' angles,
    // the runs and the jump table
    auto run_layout = [&](size_t nruns) -> size_t {
        size_t off = 4096;
        pl->off_rowinfo = (unsigned)off;
        off += align16(4 * ((size_t)n + 1));
        pl->off_col = (unsigned)off;
        off += align16(2 * (m2 + 16));
        pl->off_J = (unsigned)off;
        off += align16(8 * (m2 + 16));
        pl->off_dE = (unsigned)off;
        pl->off_ang = (unsigned)off;
        off += align16((size_t)n * 4);
        pl->off_runs = (unsigned)off;
        off += align16(4 * nruns);
        pl->off_jump = (unsigned)off;
        return off + 16 * (size_t)misa::kJumpEntries;
    };
    pl->wave_per_read = 0;
    // Sequential angle proposals and rows of at most 64 entries: one
    // wave per read (rotor_wave_kernel), four reads per block, adjacency in LDS.
    int rowmax = 0;
    for (int v = 0; v < n; ++v) rowmax = std::max(rowmax, p->rowptr[v + 1] - p->rowptr[v]);
    bool wave_ok = !ne && !pl->randomize && rowmax <= 64 && n <= 65535 && m2 >= 1;
    if (const char* f = std::getenv("MISA_ROTOR_WAVE"))  // 0: the lane-group kernel
        if (f[0] == '0') wave_ok = false;
    // Angle proposals and rows of at most 16 entries: the run kernel.
    pl->run_kernel = 0;
    bool run_ok = wave_ok && rowmax <= 16;
    if (const char* f = std::getenv("MISA_ROTOR_RUNS"))  // 0: the wave kernel
        if (f[0] == '0') run_ok = false;
    if (run_ok) {
        pl->runs = rotor_runs(p);
        if (run_layout(pl->runs.size()) <= lmax && (pl->d_jump = seg_jump_table())) {
            pl->run_kernel = 1;
            pl->wave_per_read = 1;
            pl->threads = 256;
            pl->lanes_log2 = 6;
            pl->rpb = 4;
            pl->rpb_log2 = 2;
            pl->csr_in_lds = 1;
            pl->lds_bytes = (unsigned)run_layout(pl->runs.size());
            // the member table: per run 16 slots (member j of the run, member
            // 0 past its length): row start, variable | degree << 16 | length << 24
            std::vector<uint32_t> rows;
            for (uint32_t w : pl->runs) {
                const int v0 = (int)(w & 0xffffu), K = (int)(w >> 16);
                for (int j = 0; j < 16; ++j) {
                    const int v = v0 + (j < K ? j : 0);
                    rows.push_back((uint32_t)p->rowptr[v]);
                    rows.push_back((uint32_t)v | (uint32_t)(p->rowptr[v + 1] - p->rowptr[v]) << 16 | (uint32_t)K << 24);
                }
            }
            if (int rc = dev_upload(&pl->d_runs, pl->runs)) return rc;
            return dev_upload(&pl->d_run_rows, rows);
        }
    }
    if (wave_ok && layout(1, 4) <= lmax) {
        pl->wave_per_read = 1;
        pl->threads = 256;
        pl->lanes_log2 = 6;
        pl->rpb = 4;
        pl->rpb_log2 = 2;
        pl->csr_in_lds = 1;
        pl->lds_bytes = (unsigned)layout(1, 4);
        return MISA_OK;
    }
    for (int csr = (n <= 65535 ? 1 : 0); csr >= 0; --csr) {
        int rpb = 256;
        while (rpb > 1 && layout(csr, rpb) > lmax) rpb /= 2;
        if (layout(csr, rpb) > lmax) continue;
        int L = rpb >= 64 ? 1 : 64 / rpb;
        if (ne) {
            int want = 1;
            while (want < 16 && want < p->maxdeg) want *= 2;
            if (want > L) L = want;
        }
        if (const char* f = std::getenv("MISA_ROTOR_LANES")) {  // tuning override
            const int want = std::atoi(f);
            if (want >= 1 && want <= 64 && !(want & (want - 1))) L = want;
        }
        rpb = std::max(1, std::min(rpb, 256 / L));
        pl->threads = std::max(64, ((rpb * L + 63) / 64) * 64);
        pl->lanes_log2 = log2_exact(L);
        pl->rpb = rpb;
        pl->rpb_log2 = log2_exact(rpb);
        pl->csr_in_lds = csr;
        pl->lds_bytes = (unsigned)layout(csr, rpb);
        return MISA_OK;
    }
    return fail(MISA_ERR_UNSUPPORTED,
                "problem too large for the rotor kernel's local memory (" + std::to_string(n) + " variables)");
}

void fill_args(const misa_rotor_plan* pl, misa::RotorArgs& a) {
    const misa_problem* p = pl->p;
    a.n = p->n;
    a.spb = pl->spb;
    a.nf = pl->nf;
    a.lanes_log2 = pl->lanes_log2;
    a.rpb = pl->rpb;
    a.rpb_log2 = pl->rpb_log2;
    a.threads = pl->threads;
    a.csr_in_lds = pl->csr_in_lds;
    a.wave_per_read = pl->wave_per_read;
    a.row_max = 0;
    a.row_min = p->n > 0 ? INT32_MAX : 0;
    for (int v = 0; v < p->n; ++v) {
        a.row_max = std::max(a.row_max, p->rowptr[v + 1] - p->rowptr[v]);
        a.row_min = std::min(a.row_min, p->rowptr[v + 1] - p->rowptr[v]);
    }
    a.m = p->m;
    a.m2 = 2 * p->m;
    a.h = p->d_h;
    a.tf = pl->d_tf;
    a.rowptr = p->d_rowptr;
    a.col = p->d_col;
    a.J = p->d_Jd;
    a.serial = p->d_serial;
    a.cs = p->d_cs;
    a.ce = p->d_ce;
    a.cw = p->d_cw;
    a.hp = pl->d_hp;
    a.hd = pl->d_hd;
    a.tables = pl->d_tables;
    a.off_rowinfo = pl->off_rowinfo;
    a.off_col = pl->off_col;
    a.off_J = pl->off_J;
    a.off_dE = pl->off_dE;
    a.off_ang = pl->off_ang;
    a.off_runs = pl->off_runs;
    a.off_jump = pl->off_jump;
    a.lds_bytes = pl->lds_bytes;
    a.run_kernel = pl->run_kernel;
    a.nruns = (int)pl->runs.size();
    a.runs = pl->d_runs;
    a.run_rows = pl->d_run_rows;
    a.jump = pl->d_jump;
}

}  // namespace

extern "C" {

int misa_rotor_plan_create(const misa_problem* pc, const double* trans_fields, int sweeps_per_beta, const double* hp_field,
                           const double* hd_field, int num_fields, int randomize_order, int proposal,
                           misa_rotor_plan** out) {
    if (!pc || !out) return fail(MISA_ERR_ARG, "null pointer");
    *out = nullptr;
    misa_problem* p = const_cast<misa_problem*>(pc);
    if (proposal < 0 || proposal > 3) return fail(MISA_ERR_ARG, "Unknown proposal_acceptance_criteria");
    if (sweeps_per_beta < 0 || num_fields < 0) return fail(MISA_ERR_ARG, "negative schedule length");
    if (num_fields > 0 && (!hp_field || !hd_field)) return fail(MISA_ERR_ARG, "null schedule");
    if (p->n > 0 && !trans_fields) return fail(MISA_ERR_ARG, "null trans_fields");
    int rc;
    if ((rc = p->ensure_device())) return rc;
    auto* pl = new misa_rotor_plan();
    pl->p = p;
    pl->spb = sweeps_per_beta;
    pl->nf = num_fields;
    pl->randomize = randomize_order ? 1 : 0;
    pl->proposal = proposal;
    // cpu_rotormc.cpp:392-396
    std::vector<double> tables(512);
    for (int index = 0; index < 256; ++index) {
        tables[index] = std::cos(2 * index * M_PI / 256);
        tables[256 + index] = std::abs(std::sin(2 * index * M_PI / 256));
    }
    std::vector<double> tf(trans_fields, trans_fields + p->n), hp(hp_field, hp_field + num_fields),
        hd(hd_field, hd_field + num_fields);
    if ((rc = rotor_geometry(pl)) || (rc = dev_upload(&pl->d_tf, tf)) || (rc = dev_upload(&pl->d_hp, hp)) ||
        (rc = dev_upload(&pl->d_hd, hd)) || (rc = dev_upload(&pl->d_tables, tables))) {
        std::string keep = g_err;
        delete pl;
        return fail(rc, keep);
    }
    *out = pl;
    return MISA_OK;
}

void misa_rotor_plan_destroy(misa_rotor_plan* plan) { delete plan; }

int misa_rotor_plan_info(const misa_rotor_plan* pl, misa_rotor_stats* info) {
    if (!pl || !info) return fail(MISA_ERR_ARG, "null pointer");
    std::memset(info, 0, sizeof(*info));
    info->lanes_per_read = 1 << pl->lanes_log2;
    info->reads_per_block = pl->rpb;
    info->threads_per_block = pl->threads;
    info->csr_in_lds = pl->csr_in_lds;
    info->lds_bytes = (int)pl->lds_bytes;
    info->wave_per_read = pl->wave_per_read;
    info->run_kernel = pl->run_kernel;
    return MISA_OK;
}

int misa_rotor_run_device(const misa_rotor_plan* pl, uint8_t* d_states, double* d_energies, const uint64_t* d_seeds,
                          int num_samples, uint8_t* d_statistics, int schedule_sample_interval,
                          unsigned long long* d_accepted, void* stream) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    if (schedule_sample_interval < 0) return fail(MISA_ERR_ARG, "schedule_sample_interval must be non-negative");
    if (num_samples == 0 || pl->p->n == 0) return MISA_OK;
    if (!d_states || !d_energies || !d_seeds) return fail(MISA_ERR_ARG, "null buffer");
    misa::RotorArgs a{};
    fill_args(pl, a);
    a.R = num_samples;
    a.r0 = 0;
    a.seeds = d_seeds;
    a.states = d_states;
    a.energies = d_energies;
    a.counters = d_accepted;
    a.interval = schedule_sample_interval;
    a.snapshots = (schedule_sample_interval > 0 && pl->nf > 0) ? 1 + (pl->nf - 1) / schedule_sample_interval : 0;
    a.stats = a.snapshots ? d_statistics : nullptr;
    const int nblocks = (num_samples + pl->rpb - 1) / pl->rpb;
    MISA_HIP(misa::launch_rotor(a, pl->randomize, pl->proposal, nblocks, reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

int misa_rotor_sample(const misa_problem* p, uint8_t* states, double* energies, int num_samples,
                      const double* trans_fields, int sweeps_per_beta, const double* hp_field,
                      const double* hd_field, int num_fields, const uint64_t* read_seeds, int randomize_order,
                      int proposal, uint8_t* statistics, int schedule_sample_interval, misa_interrupt_cb interrupt,
                      void* interrupt_ctx, misa_rotor_stats* stats) {
    if (!p) return fail(MISA_ERR_ARG, "null problem");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    if (proposal < 0 || proposal > 3) return fail(MISA_ERR_ARG, "Unknown proposal_acceptance_criteria");
    if (schedule_sample_interval < 0) return fail(MISA_ERR_ARG, "schedule_sample_interval must be non-negative");
    const int n = p->n;
    // rotormc_annealing.pyx:161-166: nothing to sample.
    if ((int64_t)num_samples * n == 0) {
        for (int r = 0; r < num_samples; ++r) energies[r] = 0.0;
        return num_samples;
    }
    if (!states || !energies || !read_seeds) return fail(MISA_ERR_ARG, "null buffer");
    misa_rotor_plan* pl = nullptr;
    int rc = misa_rotor_plan_create(p, trans_fields, sweeps_per_beta, hp_field, hd_field, num_fields,
                                    randomize_order, proposal, &pl);
    if (rc) return rc;
    const size_t total = (size_t)num_samples * n;
    const int snaps = (schedule_sample_interval > 0 && num_fields > 0)
                          ? 1 + (num_fields - 1) / schedule_sample_interval : 0;
    const size_t snap_read = (size_t)snaps * n;
    uint8_t* d_states = nullptr;
    uint8_t* d_stats = nullptr;
    double* d_en = nullptr;
    uint64_t* d_seeds = nullptr;
    unsigned long long* d_acc = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int done = 0;
    float total_ms = 0.f;
    auto cleanup = [&]() {
        if (d_states) (void)hipFree(d_states);
        if (d_stats) (void)hipFree(d_stats);
        if (d_en) (void)hipFree(d_en);
        if (d_seeds) (void)hipFree(d_seeds);
        if (d_acc) (void)hipFree(d_acc);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        delete pl;
    };
#define RT_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            cleanup();                                                                     \
            return fail(MISA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                  \
    } while (0)
    RT_TRY(hipMalloc(&d_states, total));
    RT_TRY(hipMalloc(&d_en, sizeof(double) * num_samples));
    RT_TRY(hipMalloc(&d_seeds, sizeof(uint64_t) * num_samples));
    RT_TRY(hipMalloc(&d_acc, sizeof(unsigned long long) * num_samples));
    if (statistics && snap_read) RT_TRY(hipMalloc(&d_stats, snap_read * num_samples));
    RT_TRY(hipMemcpy(d_states, states, total, hipMemcpyHostToDevice));
    RT_TRY(hipMemcpy(d_seeds, read_seeds, sizeof(uint64_t) * num_samples, hipMemcpyHostToDevice));
    RT_TRY(hipEventCreate(&e0));
    RT_TRY(hipEventCreate(&e1));
    // Without an interrupt callback all reads go in one launch; with one, in
    // chunks of one block per CU, polling after every read in order
    // (cpu_rotormc.cpp:556-557).
    int chunk = num_samples;
    if (interrupt) chunk = std::max(pl->rpb * device_cus(), 1);
    std::vector<unsigned long long> acc(num_samples, 0);
    for (int r0 = 0; r0 < num_samples;) {
        const int R = std::min(chunk, num_samples - r0);
        RT_TRY(hipEventRecord(e0, nullptr));
        rc = misa_rotor_run_device(pl, d_states + (size_t)r0 * n, d_en + r0, d_seeds + r0, R,
                                   d_stats ? d_stats + (size_t)r0 * snap_read : nullptr,
                                   d_stats ? schedule_sample_interval : 0, d_acc + r0, nullptr);
        if (rc) {
            std::string keep = g_err;
            cleanup();
            return fail(rc, keep);
        }
        RT_TRY(hipEventRecord(e1, nullptr));
        RT_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        RT_TRY(hipEventElapsedTime(&ms, e0, e1));
        total_ms += ms;
        RT_TRY(hipMemcpy(states + (size_t)r0 * n, d_states + (size_t)r0 * n, (size_t)R * n, hipMemcpyDeviceToHost));
        RT_TRY(hipMemcpy(energies + r0, d_en + r0, sizeof(double) * R, hipMemcpyDeviceToHost));
        RT_TRY(hipMemcpy(acc.data() + r0, d_acc + r0, sizeof(unsigned long long) * R, hipMemcpyDeviceToHost));
        if (d_stats)
            RT_TRY(hipMemcpy(statistics + (size_t)r0 * snap_read, d_stats + (size_t)r0 * snap_read,
                             (size_t)R * snap_read, hipMemcpyDeviceToHost));
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
    if (stats) {
        misa_rotor_plan_info(pl, stats);
        stats->kernel_ms = total_ms;
        for (int r = 0; r < done; ++r) stats->accepted += (int64_t)acc[r];
    }
    cleanup();
#undef RT_TRY
    return done;
}

}  // extern "C"
