// ORACLE — test infrastructure only.
//
// Thin driver around the reference steepest-descent core, compiled from the
// reference's own source file (dwave/samplers/greedy/src/descent.cpp) where it
// lies under the reference tree, by oracle/Makefile, into oracle/_ref/.  It
// calls steepest_gradient_descent (descent.cpp:337-406) as the Cython layer
// does (greedy/descent.pyx:97-101), spreading the samples over std::threads
// (one call per chunk; the solver keeps no state across samples).  Used for
// parity checks and as the reference CPU timing; never shipped.
#include <cstdint>
#include <thread>
#include <vector>

#include "descent.h"

extern "C" {

int ref_descent(int8_t* states, double* energies, uint32_t* steps, int R, int n, const double* h, int m,
                const int* cs, const int* ce, const double* cw, int large_sparse_opt, int nthreads) {
    std::vector<double> hv(h, h + n), wv(cw, cw + m);
    std::vector<int> sv(cs, cs + m), ev(ce, ce + m);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > R) nthreads = R > 0 ? R : 1;
    std::vector<int> status(nthreads, 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&, t]() {
            const int a = (int)((int64_t)R * t / nthreads), b = (int)((int64_t)R * (t + 1) / nthreads);
            if (b <= a) return;
            try {
                dwave::samplers::greedy::steepest_gradient_descent(states + (size_t)a * n, energies + a,
                                                                   reinterpret_cast<unsigned*>(steps + a), b - a,
                                                                   hv, sv, ev, wv, large_sparse_opt != 0);
            } catch (...) {
                status[t] = -1;
            }
        });
    for (auto& th : pool) th.join();
    for (int s : status)
        if (s) return -1;
    return 0;
}

}  // extern "C"
