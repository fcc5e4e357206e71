// ORACLE — test infrastructure only.
//
// Thin driver around the reference simulated-annealing core, compiled from the
// reference's own source file (dwave/samplers/sa/src/cpu_sa.cpp) where it lies
// under the reference tree, by oracle/Makefile, into oracle/_ref/.  It calls
// general_simulated_annealing (cpu_sa.h:63-78) exactly as the Cython layer
// does (sa/simulated_annealing.pyx:172-186), either once per read with
// num_samples=1 and that read's own seed (read-by-read parity), or once for
// all reads (call-level semantics).  Used to write golden vectors and as the
// "reference" CPU baseline; the product never links it.
#include <cstdint>
#include <thread>
#include <vector>

#include "cpu_sa.h"

using dwave::samplers::sa::general_simulated_annealing;
using dwave::samplers::sa::Gibbs;
using dwave::samplers::sa::Metropolis;
using dwave::samplers::sa::Random;
using dwave::samplers::sa::Sequential;

static bool never_interrupt(void* const) { return false; }

extern "C" {

// One call covering all R reads (one RNG stream).  Returns the completed
// count, or -1 if the reference threw (mismatched lengths / bad index).
int ref_sa_call(int8_t* states, double* energies, int R, int n, const double* h,
                int m, const int* cs, const int* ce, const double* cw,
                int sweeps_per_beta, const double* betas, int nbetas,
                uint64_t seed, int order, int proposal) {
    try {
        return general_simulated_annealing(
            states, energies, R, std::vector<double>(h, h + n),
            std::vector<int>(cs, cs + m), std::vector<int>(ce, ce + m),
            std::vector<double>(cw, cw + m), sweeps_per_beta,
            std::vector<double>(betas, betas + nbetas), seed,
            order ? Random : Sequential, proposal ? Metropolis : Gibbs,
            never_interrupt, nullptr);
    } catch (...) {
        return -1;
    }
}

// Read-by-read: read r is its own call with num_samples=1 and seeds[r],
// spread over nthreads std::threads (the reference RNG is thread_local,
// cpu_sa.cpp:41, so concurrent calls are independent).
int ref_sa_read_by_read(int8_t* states, double* energies, int R, int n,
                        const double* h, int m, const int* cs, const int* ce,
                        const double* cw, int sweeps_per_beta, const double* betas,
                        int nbetas, const uint64_t* seeds, int order, int proposal,
                        int nthreads) {
    std::vector<double> hv(h, h + n), wv(cw, cw + m), bv(betas, betas + nbetas);
    std::vector<int> sv(cs, cs + m), ev(ce, ce + m);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R > 0 ? R : 1;
    std::vector<int> status(nthreads, 0);
    auto work = [&](int t) {
        for (int r = (int)((int64_t)R * t / nthreads);
             r < (int)((int64_t)R * (t + 1) / nthreads); ++r) {
            try {
                general_simulated_annealing(
                    states + (int64_t)r * n, energies + r, 1, hv, sv, ev, wv,
                    sweeps_per_beta, bv, seeds[r], order ? Random : Sequential,
                    proposal ? Metropolis : Gibbs, never_interrupt, nullptr);
            } catch (...) {
                status[t] = -1;
                return;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    for (int t = 0; t < nthreads; ++t)
        if (status[t]) return -1;
    return R;
}

}  // extern "C"
