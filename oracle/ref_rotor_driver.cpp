// ORACLE — test infrastructure only.
//
// Thin driver around the reference rotor-model core, compiled from the
// reference's own source file (dwave/samplers/sqa/src/cpu_rotormc.cpp) where
// it lies under the reference tree, by oracle/Makefile, into oracle/_ref/.  It
// calls general_simulated_annealing (cpu_rotormc.cpp:366-562) once per read
// with num_samples = 1 and that read's seed, as the engine defines a read,
// spreading reads over std::threads (the core's RNG state is thread_local).
// Snapshots go to (read, 1 + (len - 1) / interval, n), the Cython layer's
// shape (rotormc_annealing.pyx:171-176).  Used for parity checks and as the
// reference CPU timing; never shipped.
#include <cstdint>
#include <thread>
#include <vector>

#include "cpu_rotormc.h"

namespace {
bool never_interrupt(void* const) { return false; }
}  // namespace

extern "C" {

int ref_rotor_call(uint8_t* states, double* energies, int R, int n, const double* h, int m, const int* cs,
                   const int* ce, const double* cw, const double* tf, int spb, const double* hp,
                   const double* hd, int nf, const uint64_t* seeds, int randomize, int proposal,
                   uint8_t* stats, int interval, int nthreads) {
    using namespace dwave::samplers::sqa;
    std::vector<double> hv(h, h + n), wv(cw, cw + m), tv(tf, tf + n), hpv(hp, hp + nf), hdv(hd, hd + nf);
    std::vector<int> sv(cs, cs + m), ev(ce, ce + m);
    const int64_t block = (interval > 0 && nf > 0) ? (int64_t)n * (1 + (nf - 1) / interval) : 0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R > 0 ? R : 1;
    std::vector<int> status(nthreads, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&, t]() {
            const int a = (int)((int64_t)R * t / nthreads), b = (int)((int64_t)R * (t + 1) / nthreads);
            std::vector<uint8_t> scratch(1);
            for (int r = a; r < b; ++r) {
                try {
                    uint8_t* st = stats && block ? stats + (size_t)r * block : scratch.data();
                    general_simulated_annealing(states + (size_t)r * n, energies + r, 1, hv, sv, ev, wv, tv, spb,
                                                hpv, hdv, seeds[r], randomize != 0, (Proposal)proposal, st,
                                                stats ? interval : 0, never_interrupt, nullptr);
                } catch (...) {
                    status[t] = -1;
                    return;
                }
            }
        });
    for (auto& th : pool) th.join();
    for (int s : status)
        if (s) return -1;
    return 0;
}

}  // extern "C"
