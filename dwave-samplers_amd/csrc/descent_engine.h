// Steepest-descent kernel launch arguments (descent_api.cpp <-> descent_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

struct DescentArgs {
    int n;            // variables
    int R;            // reads in this launch
    int waves;        // reads (waves) per block
    int wave_lds;     // LDS bytes per wave: dE (8n) + chunk minima (12 ceil(n/64)) + spins (n)
    uint32_t max_steps;
    int64_t m;        // couplers in input order (final energy)
    const double* h;
    const int* rowptr;    // adjacency in coupler insertion order
    const int* col;
    const double* J;
    const uint8_t* serial;  // row repeats a neighbour or holds the variable itself
    const int* cs;
    const int* ce;
    const double* cw;
    int8_t* states;       // R x n in/out
    double* energies;     // R
    uint32_t* steps;      // R
};

hipError_t launch_descent(const DescentArgs& a, hipStream_t stream);

}  // namespace misa
