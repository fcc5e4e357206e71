// C-ABI host side of the steepest-descent solver (include/misamplers.h).
//
// Replaces the reference's steepest_gradient_descent
// (dwave/samplers/greedy/src/descent.cpp:337-406, declared in descent.h): the
// coupler validation and adjacency are the problem object's (same insertion
// order and messages); one wave descends one read on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "descent_engine.h"
#include "host.h"
#include "misamplers.h"
#include "sa_engine.h"

using namespace misa_host;

namespace {

int descent_geometry(int n, int& waves, int& wave_lds) {
    const size_t nc = ((size_t)n + 63) / 64;
    const size_t per = ((8 * (size_t)n + 16 * nc + (size_t)n) + 15) & ~(size_t)15;
    const size_t lmax = misa::max_dynamic_lds();
    if (per > lmax)
        return fail(MISA_ERR_UNSUPPORTED, "problem too large for the descent kernel's local memory (" +
                                              std::to_string(n) + " variables)");
    waves = (int)std::min<size_t>(4, lmax / per);
    wave_lds = (int)per;
    return MISA_OK;
}

// Padded rows for the kernel (descent_engine.h): each variable's neighbours in
// coupler insertion order (descent.cpp:366-380) without self-loops, so that a
// step's row loads depend only on the chosen variable.  Used when every row
// fits one pass of the wave and padding at most quadruples the entries.
int exact_sum_tables(misa_problem* p);

int ensure_rows(misa_problem* p) {
    std::lock_guard<std::mutex> lk(p->mu);
    if (p->ell_width >= 0) return MISA_OK;
    const int n = p->n;
    std::vector<int> deg(n, 0);
    int width = 0;
    int64_t entries = 0;
    for (int v = 0; v < n; ++v) {
        for (int k = p->rowptr[v]; k < p->rowptr[v + 1]; ++k)
            if (p->col[k] != v) ++deg[v];
        width = std::max(width, deg[v]);
        entries += deg[v];
    }
    width = std::max(width, 1);
    if (width > misa::DESCENT_MAX_WIDTH || n > misa::DESCENT_COL_MASK ||
        (int64_t)n * width > 4 * std::max<int64_t>(entries, n)) {
        p->ell_width = 0;
        return exact_sum_tables(p);
    }
    std::vector<int> col((size_t)n * width, misa::DESCENT_PAD);
    std::vector<double> J((size_t)n * width, 0.0);
    std::vector<int> seen(n, -1);
    for (int v = 0; v < n; ++v) {
        bool repeats = false;
        int j = 0;
        for (int k = p->rowptr[v]; k < p->rowptr[v + 1]; ++k) {
            const int u = p->col[k];
            if (u == v) continue;
            if (seen[u] == v) repeats = true;
            seen[u] = v;
            col[(size_t)v * width + j] = u;
            J[(size_t)v * width + j] = p->Jd[k];
            ++j;
        }
        if (repeats)
            for (int i = 0; i < j; ++i) col[(size_t)v * width + i] |= misa::DESCENT_ROW_SERIAL;
    }
    int rc;
    if ((rc = dev_upload(&p->d_ell_col, col)) || (rc = dev_upload(&p->d_ell_J, J))) {
        dev_free(p->d_ell_col);
        dev_free(p->d_ell_J);
        return rc;
    }
    p->ell_width = width;
    return exact_sum_tables(p);
}

// Whether every energy sum is exact (see misa_problem::exact_sum): the least
// power of two dividing every nonzero h and J, and sum |h| + sum |J| against
// 2^52 of it.  Then the kernel adds integer multiples in any order.
int exact_sum_tables(misa_problem* p) {
    int lo = INT32_MAX;
    long double total = 0.0L;
    auto scan = [&](double x) {
        if (!std::isfinite(x)) return false;
        if (x == 0.0) return true;
        int e;
        const double f = std::frexp(std::fabs(x), &e);   // |x| = f 2^e, f in [0.5, 1)
        uint64_t mant = (uint64_t)std::ldexp(f, 53);      // exact: 53-bit integer
        int low = e - 53;
        while (!(mant & 1)) {
            mant >>= 1;
            ++low;
        }
        lo = std::min(lo, low);
        total += std::fabs((long double)x);
        return true;
    };
    for (double x : p->h)
        if (!scan(x)) return MISA_OK;
    for (double x : p->cw)
        if (!scan(x)) return MISA_OK;
    if (lo == INT32_MAX) lo = 0;  // all zero
    if (!(std::ldexp(total, -lo) < std::ldexp(1.0L, 52))) return MISA_OK;
    std::vector<int64_t> hq(p->h.size()), wq(p->cw.size());
    for (size_t i = 0; i < hq.size(); ++i) hq[i] = (int64_t)std::ldexp(p->h[i], -lo);
    for (size_t i = 0; i < wq.size(); ++i) wq[i] = (int64_t)std::ldexp(p->cw[i], -lo);
    int rc;
    if ((rc = dev_upload(&p->d_h_q, hq)) || (rc = dev_upload(&p->d_cw_q, wq))) {
        dev_free(p->d_h_q);
        dev_free(p->d_cw_q);
        return rc;
    }
    p->exact_exp = lo;
    p->exact_sum = true;
    return MISA_OK;
}

}  // namespace

extern "C" {

int misa_descent_run_device(const misa_problem* pc, int8_t* d_states, double* d_energies, uint32_t* d_steps,
                            int num_samples, uint32_t max_steps, void* stream) {
    if (!pc) return fail(MISA_ERR_ARG, "null problem");
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    if (num_samples == 0 || pc->n == 0) return MISA_OK;
    if (!d_states || !d_energies || !d_steps) return fail(MISA_ERR_ARG, "null buffer");
    misa_problem* p = const_cast<misa_problem*>(pc);
    int rc = p->ensure_device();
    if (rc || (rc = ensure_rows(p))) return rc;
    misa::DescentArgs a{};
    if ((rc = descent_geometry(p->n, a.waves, a.wave_lds))) return rc;
    a.n = p->n;
    a.R = num_samples;
    a.max_steps = max_steps;
    a.m = p->m;
    a.h = p->d_h;
    a.ell_width = p->ell_width;
    a.ell_col = p->d_ell_col;
    a.ell_J = p->d_ell_J;
    a.rowptr = p->d_rowptr;
    a.col = p->d_col;
    a.J = p->d_Jd;
    a.serial = p->d_serial;
    a.cs = p->d_cs;
    a.ce = p->d_ce;
    a.cw = p->d_cw;
    a.exact_sum = p->exact_sum ? 1 : 0;
    a.exact_exp = p->exact_exp;
    a.h_q = p->d_h_q;
    a.cw_q = p->d_cw_q;
    a.states = d_states;
    a.energies = d_energies;
    a.steps = d_steps;
    MISA_HIP(misa::launch_descent(a, reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

int misa_descent_sample(const misa_problem* p, int8_t* states, double* energies, uint32_t* num_steps,
                        int num_samples, uint32_t max_steps, double* kernel_ms) {
    if (!p) return fail(MISA_ERR_ARG, "null problem");
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    if (kernel_ms) *kernel_ms = 0.0;
    const int n = p->n;
    // greedy/descent.pyx:86-91: nothing to descend.
    if ((int64_t)num_samples * n == 0) {
        for (int r = 0; r < num_samples; ++r) {
            energies[r] = 0.0;
            num_steps[r] = 0;
        }
        return MISA_OK;
    }
    if (!states || !energies || !num_steps) return fail(MISA_ERR_ARG, "null buffer");
    const size_t total = (size_t)num_samples * n;
    for (size_t i = 0; i < total; ++i)
        if (states[i] != 1 && states[i] != -1) return fail(MISA_ERR_ARG, "initial states must be +1/-1 spins");
    int waves = 0, wl = 0;
    int rc = descent_geometry(n, waves, wl);
    if (rc) return rc;
    int8_t* d_s = nullptr;
    double* d_e = nullptr;
    uint32_t* d_k = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (d_s) (void)hipFree(d_s);
        if (d_e) (void)hipFree(d_e);
        if (d_k) (void)hipFree(d_k);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
#define GD_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            cleanup();                                                                     \
            return fail(MISA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                  \
    } while (0)
    GD_TRY(hipMalloc(&d_s, total));
    GD_TRY(hipMalloc(&d_e, sizeof(double) * num_samples));
    GD_TRY(hipMalloc(&d_k, sizeof(uint32_t) * num_samples));
    GD_TRY(hipMemcpy(d_s, states, total, hipMemcpyHostToDevice));
    GD_TRY(hipEventCreate(&e0));
    GD_TRY(hipEventCreate(&e1));
    GD_TRY(hipEventRecord(e0, nullptr));
    rc = misa_descent_run_device(p, d_s, d_e, d_k, num_samples, max_steps, nullptr);
    if (rc) {
        std::string keep = g_err;
        cleanup();
        return fail(rc, keep);
    }
    GD_TRY(hipEventRecord(e1, nullptr));
    GD_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    GD_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (kernel_ms) *kernel_ms = ms;
    GD_TRY(hipMemcpy(states, d_s, total, hipMemcpyDeviceToHost));
    GD_TRY(hipMemcpy(energies, d_e, sizeof(double) * num_samples, hipMemcpyDeviceToHost));
    GD_TRY(hipMemcpy(num_steps, d_k, sizeof(uint32_t) * num_samples, hipMemcpyDeviceToHost));
    cleanup();
#undef GD_TRY
    return MISA_OK;
}

}  // extern "C"
