// Steepest-descent kernel launch arguments (descent_api.cpp <-> descent_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

// Padded-row entries: the neighbour index, with DESCENT_ROW_SERIAL set on every
// entry of a row that names one neighbour twice (its updates must be applied in
// order by one lane); DESCENT_PAD fills a row past its degree.
constexpr int DESCENT_ROW_SERIAL = 1 << 30;
constexpr int DESCENT_COL_MASK = DESCENT_ROW_SERIAL - 1;
constexpr int DESCENT_PAD = -1;
constexpr int DESCENT_MAX_WIDTH = 64;

// LDS bytes per wave of the energy kernel: 128 doubles of staging + the spins.
__host__ __device__ inline size_t descent_energy_wave_lds(int n) { return 1024 + (((size_t)n + 15) & ~(size_t)15); }

struct DescentArgs {
    int n;            // variables
    int R;            // reads in this launch
    int waves;        // reads (waves) per block
    int wave_lds;     // LDS bytes per wave: dE (8n) + chunk slots (16 ceil(n/64)) + spins (n)
    uint32_t max_steps;
    int64_t m;        // couplers in input order (final energy)
    const double* h;
    // Rows in coupler insertion order without self-loops.  ell_width > 0:
    // row v is ell_col / ell_J [v * ell_width, (v + 1) * ell_width).
    // ell_width == 0: the CSR (rowptr, col, J, serial), whose self entries the
    // kernel skips.
    int ell_width;
    const int* ell_col;
    const double* ell_J;
    const int* rowptr;
    const int* col;
    const double* J;
    const uint8_t* serial;  // CSR row repeats a neighbour or holds the variable itself
    const int* cs;
    const int* ce;
    const double* cw;
    // Exact energies (sum order free): h = h_q 2^exact_exp, J = cw_q 2^exact_exp.
    int exact_sum;
    int exact_exp;
    const int64_t* h_q;
    const int64_t* cw_q;
    int8_t* states;       // R x n in/out
    double* energies;     // R
    uint32_t* steps;      // R
};

// The descent (states, step counts) followed by the final energies.
hipError_t launch_descent(const DescentArgs& a, hipStream_t stream);

}  // namespace misa
