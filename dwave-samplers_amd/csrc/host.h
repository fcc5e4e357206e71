// Host-side pieces shared by the engine's C-ABI translation units: the
// per-thread error text, HIP error mapping, device upload helpers and the
// problem object (couplers plus adjacency in coupler insertion order).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "misamplers.h"
#include "sa_engine.h"

namespace misa_host {

extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define MISA_HIP(expr)                                                             \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess)                                                      \
            return misa_host::fail(MISA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dev_upload(T** dst, const std::vector<T>& src) {
    *dst = nullptr;
    size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
    MISA_HIP(hipMalloc(reinterpret_cast<void**>(dst), bytes));
    if (!src.empty()) MISA_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return MISA_OK;
}

// As dev_upload, followed by `pad` zero elements (readable spare entries).
template <typename T>
int dev_upload_padded(T** dst, const std::vector<T>& src, size_t pad) {
    *dst = nullptr;
    const size_t count = src.size() + pad;
    MISA_HIP(hipMalloc(reinterpret_cast<void**>(dst), std::max<size_t>(count, 1) * sizeof(T)));
    if (!src.empty()) MISA_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    if (pad) MISA_HIP(hipMemset(*dst + src.size(), 0, pad * sizeof(T)));
    return MISA_OK;
}

template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int device_cus();

// The 64-step xorshift128+ jump table (kJumpEntries uint4) of the run
// kernels' per-lane streams on the current device, uploaded once per device
// and held until the process ends (null on failure).
const uint32_t* seg_jump_table();
uint64_t splitmix_at(uint64_t seed, int64_t r);

// misa_sa_run_device; with ext_state (ext_bytes of device memory) the general
// kernel's per-read state goes there instead of a plan-owned buffer, and the
// call allocates nothing.
int sa_run_device_impl(misa_sa_plan* plan, int8_t* d_states, double* d_energies, const uint64_t* d_seeds,
                       int num_samples, uint64_t* d_counters, void* stream, void* ext_state, size_t ext_bytes);

}  // namespace misa_host

struct misa_problem {
    int n = 0;
    int64_t m = 0;
    std::vector<double> h, cw;
    std::vector<int> cs, ce;
    // adjacency, cpu_sa.cpp:298-316 order
    std::vector<int> rowptr, col;
    std::vector<double> Jd;
    std::vector<int16_t> Ji;
    std::vector<uint8_t> serial;
    std::vector<int> jnext;  // per v: sum of J over row-v entries whose column is (v+1) mod n
    std::vector<double> jnext_d;   // per v: J(v, (v+1) mod n) when one row-v entry holds it, else 0
    std::vector<uint8_t> next_adj;  // per v: 1 when exactly one row-v entry has column (v+1) mod n
    std::vector<uint32_t> rowinfo;  // (rowptr[v] << 10) | (jnext[v] & 0x3ff), n+1 entries
    bool integer_valued = false;
    double fmax = 0.0;  // max_v |h_v| + sum_k |J_k|
    int maxdeg = 0;
    int reg_deg = -1;  // every row has this many entries, or -1
    // device copies
    std::mutex mu;
    bool on_device = false;
    int *d_rowptr = nullptr, *d_col = nullptr, *d_cs = nullptr, *d_ce = nullptr;
    double *d_Jd = nullptr, *d_h = nullptr, *d_cw = nullptr;
    int16_t* d_Ji = nullptr;
    uint8_t* d_serial = nullptr;
    int* d_jnext = nullptr;
    double* d_jnext_d = nullptr;
    uint8_t* d_next_adj = nullptr;
    uint32_t* d_rowinfo = nullptr;
    // Steepest-descent rows (descent_api.cpp): self-loops dropped, padded to a
    // fixed width; built on first use.  ell_width < 0: not built yet, 0: rows
    // too wide or irregular for the padded form (the kernel reads the CSR).
    int ell_width = -1;
    int* d_ell_col = nullptr;
    double* d_ell_J = nullptr;
    // Exact energies: every h and J an integer multiple of 2^exact_exp with
    // sum |h| + sum |J| < 2^52 * 2^exact_exp, so that every partial sum of an
    // energy is exact and the sum is the same in any order.  h_q / cw_q are the
    // integer multiples (exact_sum: set only then).
    bool exact_sum = false;
    int exact_exp = 0;
    int64_t* d_h_q = nullptr;
    int64_t* d_cw_q = nullptr;
    // Runs for the segment kernel, one cut per run-length cap: consecutive
    // variables with no coupler between any two of them; a variable whose row
    // is serial (repeated neighbour or self-loop) is a run of its own.
    struct Runs {
        int cap = 0;
        int nseg = 0;
        int* d_starts = nullptr;  // nseg + 1 run starts (the last is n), bit 31: serial run
        // Run tiles for one read per wave (cap 32, rows of at most 64
        // entries): each member's row padded to P = 1 << tile_plog2 entries
        // (the smallest power of two >= the longest row, at least 16), 64 / P
        // members per 64-entry slot, ceil(K / (64 / P)) slots per run of K
        // members, run r at entry tile_off[4 r] (tile_off[4 r + 1], [4 r + 2]:
        // the row bounds of the run's first member, the row of a serial run);
        // a padding entry has column -1 and weight 0.
        int tile_plog2 = -1;  // -1: no tiles
        int* d_tile_off = nullptr;
        int* d_tile_col = nullptr;
        double* d_tile_J = nullptr;
    };
    std::vector<Runs> runs;
    int* d_rowptr_pad = nullptr;  // rowptr + 64 copies of rowptr[n]

    // The runs for this cap (built and uploaded on first use; n <= 65535).
    int get_runs(int cap, Runs* out) {
        using namespace misa_host;
        std::lock_guard<std::mutex> lk(mu);
        for (const Runs& r : runs)
            if (r.cap == cap) {
                *out = r;
                return MISA_OK;
            }
        if (n > 65535 || n < 1) return fail(MISA_ERR_UNSUPPORTED, "runs need 1 <= n <= 65535");
        if (!d_rowptr_pad) {
            std::vector<int> rp(rowptr);
            rp.insert(rp.end(), 64, rowptr[n]);
            int rc = dev_upload(&d_rowptr_pad, rp);
            if (rc) return rc;
        }
        std::vector<int> st;
        int start = 0;
        st.push_back(serial[0] ? (int)0x80000000u : 0);
        for (int v = 1; v < n; ++v) {
            bool join = v - start < cap && !serial[start] && !serial[v];
            for (int k = rowptr[v]; join && k < rowptr[v + 1]; ++k)
                if (col[k] >= start && col[k] < v) join = false;
            if (!join) {
                st.push_back(v | (serial[v] ? (int)0x80000000u : 0));
                start = v;
            }
        }
        st.push_back(n);
        Runs r;
        r.cap = cap;
        r.nseg = (int)st.size() - 1;
        int rc = dev_upload(&r.d_starts, st);
        if (rc) return rc;
        if (cap == 32 && maxdeg <= 64) {
            int pl = 4;
            while ((1 << pl) < maxdeg) ++pl;
            const int P = 1 << pl, G = 64 / P;
            std::vector<int> off(4 * ((size_t)r.nseg + 1), 0), tcol;
            std::vector<double> tJ;
            for (int q = 0; q < r.nseg; ++q) {
                const int v0 = st[q] & 0x7fffffff, v1 = st[q + 1] & 0x7fffffff;
                const int slots = (v1 - v0 + G - 1) / G;
                off[4 * q] = (int)tcol.size();
                off[4 * q + 1] = rowptr[v0];
                off[4 * q + 2] = rowptr[v0 + 1];
                tcol.resize(tcol.size() + (size_t)slots * 64, -1);
                tJ.resize(tJ.size() + (size_t)slots * 64, 0.0);
                for (int t = 0; t < v1 - v0; ++t) {
                    const size_t base = (size_t)off[4 * q] + (size_t)(t / G) * 64 + (size_t)(t % G) * P;
                    for (int k = rowptr[v0 + t]; k < rowptr[v0 + t + 1]; ++k) {
                        tcol[base + (k - rowptr[v0 + t])] = col[k];
                        tJ[base + (k - rowptr[v0 + t])] = Jd[k];
                    }
                }
            }
            off[4 * (size_t)r.nseg] = (int)tcol.size();
            tcol.resize(tcol.size() + 64, -1);  // a spare slot: loads past the last run stay in bounds
            tJ.resize(tJ.size() + 64, 0.0);
            if ((rc = dev_upload(&r.d_tile_off, off)) || (rc = dev_upload(&r.d_tile_col, tcol)) ||
                (rc = dev_upload(&r.d_tile_J, tJ)))
                return rc;
            r.tile_plog2 = pl;
        }
        runs.push_back(r);
        *out = r;
        return MISA_OK;
    }

    // Chain-run kernel data (int8 problems with rows of at most 7 entries and
    // n <= 65535): per variable kChainRecDwords dwords (the layout in
    // sa_engine.h) and the run starts: greedy runs of at most kChainK
    // consecutive variables in which no variable has a CSR entry naming an
    // earlier member other than its predecessor.  Built on first use.
    bool chain_tried = false;
    int chain_nruns = 0;
    int chain_ncrit = 0;  // most entries of a record naming a variable after it, up to the next run's end
    // chain_freach[f + 64]: 1 when some variable can hold field f (a superset:
    // h_v plus every sign pattern over its row's entries)
    std::vector<uint8_t> chain_freach;
    uint32_t* d_chain_rec = nullptr;
    int* d_chain_runs = nullptr;
    // Run-major slot records (rows of at most 6 entries, else null): slot
    // (run r, lane k) at 8 (16 r + k) holds member k's record with dword 7 =
    // its word row's byte offset 16 v, and bit 31 set for a lane past the run
    // (which holds the run's first member's record and adds zeros).
    uint32_t* d_chain_slots = nullptr;

    // Builds and uploads the chain-run data once; returns MISA_OK with
    // *nruns = 0 when the problem does not qualify.
    int get_chain(int* nruns) {
        using namespace misa_host;
        std::lock_guard<std::mutex> lk(mu);
        if (!chain_tried) {
            chain_tried = true;
            if (n >= 2 && n <= 65535 && maxdeg <= misa::kChainRecDwords - 1 && integer_valued && fmax <= 63.0 &&
                (int64_t)Ji.size() == 2 * m) {
                std::vector<int> st{0};
                int start = 0;
                for (int v = 1; v < n; ++v) {
                    bool join = v - start < misa::kChainK;
                    for (int k = rowptr[v]; join && k < rowptr[v + 1]; ++k)
                        if (col[k] >= start && col[k] < v - 1) join = false;
                    if (!join) {
                        st.push_back(v);
                        start = v;
                    }
                }
                st.push_back(n);
                const int nr = (int)st.size() - 1;
                // end of the run after v's run (n past the last two runs)
                std::vector<int> next_end(n);
                for (int r = 0; r < nr; ++r)
                    for (int v = st[r]; v < st[r + 1]; ++v) next_end[v] = st[std::min(r + 2, nr)];
                // records; entries that name a variable of the rest of v's run
                // or of the next run come first (the kernel issues their atomics
                // before it loads the next run's words)
                std::vector<uint32_t> rec((size_t)n * misa::kChainRecDwords, 0u);
                chain_ncrit = 0;
                for (int v = 0; v < n; ++v) {
                    uint32_t* r = &rec[(size_t)v * misa::kChainRecDwords];
                    const int kb = rowptr[v], ke = rowptr[v + 1];
                    int j1 = 0, slot = 1, ncrit = 0;
                    for (int pass = 0; pass < 2; ++pass)
                        for (int k = kb; k < ke; ++k) {
                            const int u = col[k], J = Ji[k];
                            const bool crit = u > v && u < next_end[v];
                            if (crit != (pass == 0)) continue;
                            r[slot++] = ((uint32_t)(uint8_t)(int8_t)(-J) << 24) | ((uint32_t)u << 4);
                            if (pass == 0) ++ncrit;
                            if (u == v + 1) j1 += J;
                        }
                    for (; slot < misa::kChainRecDwords; ++slot) r[slot] = (uint32_t)v << 4;
                    r[0] = ((uint32_t)(ke - kb) << 16) | (uint32_t)(uint16_t)(int16_t)(-4 * j1);
                    chain_ncrit = std::max(chain_ncrit, ncrit);
                }
                chain_freach.assign(129, 0);
                for (int v = 0; v < n; ++v) {
                    const int kb = rowptr[v], deg = rowptr[v + 1] - kb;
                    for (int sgn = 0; sgn < (1 << deg); ++sgn) {
                        int f = (int)h[v];
                        for (int j = 0; j < deg; ++j) f += (sgn >> j) & 1 ? -Ji[kb + j] : Ji[kb + j];
                        chain_freach[f + 64] = 1;  // |f| <= fmax <= 63
                    }
                }
                int rc;
                if ((rc = dev_upload(&d_chain_rec, rec)) || (rc = dev_upload(&d_chain_runs, st))) return rc;
                if (maxdeg <= misa::kChainRecDwords - 2) {
                    std::vector<uint32_t> slots((size_t)nr * 16 * misa::kChainRecDwords);
                    for (int r = 0; r < nr; ++r)
                        for (int k = 0; k < 16; ++k) {
                            const bool mem = k < st[r + 1] - st[r];
                            const int v = mem ? st[r] + k : st[r];
                            uint32_t* d = &slots[((size_t)r * 16 + k) * misa::kChainRecDwords];
                            std::copy_n(&rec[(size_t)v * misa::kChainRecDwords], misa::kChainRecDwords, d);
                            d[misa::kChainRecDwords - 1] = ((uint32_t)v << 4) | (mem ? 0u : 0x80000000u);
                        }
                    if ((rc = dev_upload(&d_chain_slots, slots))) return rc;
                }
                chain_nruns = nr;
            }
        }
        *nruns = chain_nruns;
        return MISA_OK;
    }

    // Dense coupling matrix for the dense kernel (n <= 128, no serial row):
    // n rows of 64 dense_nr doubles, J(v, u) at [v][u].  Built on first use;
    // dense_nr = 0 when the problem does not qualify.
    bool dense_tried = false;
    int dense_nr = 0;
    double* d_dense = nullptr;
    int get_dense(int* nr) {
        using namespace misa_host;
        std::lock_guard<std::mutex> lk(mu);
        if (!dense_tried) {
            dense_tried = true;
            bool ok = n >= 2 && n <= 128;
            for (int v = 0; ok && v < n; ++v) ok = !serial[v];
            if (ok) {
                const int R = n <= 64 ? 1 : 2, st = 64 * R;
                std::vector<double> d((size_t)n * st, 0.0);
                for (int v = 0; v < n; ++v)
                    for (int k = rowptr[v]; k < rowptr[v + 1]; ++k) d[(size_t)v * st + col[k]] = Jd[k];
                if (int rc = dev_upload(&d_dense, d)) return rc;
                dense_nr = R;
            }
        }
        *nr = dense_nr;
        return MISA_OK;
    }

    int ensure_device() {
        using namespace misa_host;
        std::lock_guard<std::mutex> lk(mu);
        if (on_device) return MISA_OK;
        int rc;
        // col, Jd and Ji carry 64 spare entries: lanes past a row's end read them
        if ((rc = dev_upload(&d_rowptr, rowptr)) || (rc = dev_upload_padded(&d_col, col, 64)) ||
            (rc = dev_upload_padded(&d_Jd, Jd, 64)) || (rc = dev_upload_padded(&d_Ji, Ji, 64)) ||
            (rc = dev_upload(&d_serial, serial)) || (rc = dev_upload(&d_h, h)) ||
            (rc = dev_upload(&d_cs, cs)) || (rc = dev_upload(&d_ce, ce)) || (rc = dev_upload(&d_cw, cw)) ||
            (rc = dev_upload(&d_jnext, jnext)) || (rc = dev_upload(&d_rowinfo, rowinfo)) ||
            (rc = dev_upload_padded(&d_jnext_d, jnext_d, 64)) || (rc = dev_upload(&d_next_adj, next_adj)))
            return rc;
        on_device = true;
        return MISA_OK;
    }
    ~misa_problem() {
        using misa_host::dev_free;
        dev_free(d_rowptr); dev_free(d_col); dev_free(d_Jd); dev_free(d_Ji); dev_free(d_serial);
        dev_free(d_h); dev_free(d_cs); dev_free(d_ce); dev_free(d_cw); dev_free(d_jnext);
        dev_free(d_jnext_d); dev_free(d_next_adj);
        dev_free(d_rowinfo); dev_free(d_ell_col); dev_free(d_ell_J); dev_free(d_h_q); dev_free(d_cw_q);
        for (Runs& r : runs) {
            dev_free(r.d_starts);
            dev_free(r.d_tile_off);
            dev_free(r.d_tile_col);
            dev_free(r.d_tile_J);
        }
        dev_free(d_chain_rec); dev_free(d_chain_runs); dev_free(d_chain_slots);
        dev_free(d_rowptr_pad);
        dev_free(d_dense);
    }
};
