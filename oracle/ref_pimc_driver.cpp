// ORACLE — test infrastructure only.
//
// Thin driver around the reference path-integral core, compiled from the
// reference's own source file (dwave/samplers/sqa/src/localPIMC.cpp) where it
// lies under the reference tree, by oracle/Makefile, into oracle/_ref/.  It
// calls general_simulated_annealing (localPIMC.h:60-86; definition
// localPIMC.cpp:604-715) the way the Cython layer does
// (sqa/pimc_annealing.pyx:253-278): Hp at the 15th position and Hd at the
// 16th, classical (projected) states in and out.  Reads are split into chunks,
// one call and one 32-bit seed per chunk, spread over std::threads (each call
// owns its localPIMC object and mt19937, localPIMC.cpp:680-684).  Used for
// statistical comparisons and as the reference CPU timing; never shipped.
#include <cstdint>
#include <thread>
#include <vector>

#include "localPIMC.h"

using dwave::samplers::sqa::general_simulated_annealing;

static bool never_interrupt(void* const) { return false; }

extern "C" {

// states (R x n) in/out, energies (R), one seed per chunk (nchunks seeds, read
// chunk t covers reads [R*t/nchunks, R*(t+1)/nchunks)).  stats may be null when
// interval == 0; otherwise it is (R, 1 + (nsteps-1)/interval, n).  Returns 0,
// or -1 if a reference call threw.
int ref_pimc_call(int8_t* states, double* energies, int R, int n, const double* h, int m, const int* cs,
                  const int* ce, const double* cw, int sweeps_per_beta, const double* hp, const double* hd,
                  int nsteps, const uint32_t* seeds, int nchunks, int8_t* stats, int interval, int nthreads) {
    std::vector<double> hv(h, h + n), wv(cw, cw + m), hpv(hp, hp + nsteps), hdv(hd, hd + nsteps);
    std::vector<int> sv(cs, cs + m), ev(ce, ce + m);
    if (nchunks < 1) nchunks = 1;
    if (nthreads < 1) nthreads = 1;
    const int per_read_stats = interval ? (1 + (nsteps - 1) / interval) * n : 0;
    std::vector<int> status(nchunks, 0);
    auto run_chunk = [&](int t) {
        const int a = (int)((int64_t)R * t / nchunks), b = (int)((int64_t)R * (t + 1) / nchunks);
        if (b <= a) return;
        std::vector<int> nb((size_t)(b - a) * n, 0);
        int dummy_in = 0, dummy_out = 0;
        std::vector<int8_t> dummy_stats(1, 0);
        try {
            int got = general_simulated_annealing(
                states + (size_t)a * n, energies + a, true, true, nb.data(), &dummy_in, &dummy_out, 1, b - a, hv,
                sv, ev, wv, sweeps_per_beta, hpv, hdv, 1.0, 1.0, 1, 1, seeds[t],
                interval ? stats + (size_t)a * per_read_stats : dummy_stats.data(), interval, never_interrupt,
                nullptr);
            status[t] = got == b - a ? 0 : -1;
        } catch (...) {
            status[t] = -1;
        }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < nthreads; ++w)
        pool.emplace_back([&, w]() {
            for (int t = w; t < nchunks; t += nthreads) run_chunk(t);
        });
    for (auto& th : pool) th.join();
    for (int s : status)
        if (s) return -1;
    return 0;
}

}  // extern "C"
