// Rotor-model annealing kernel launch arguments (rotor_api.cpp <-> rotor_kernel.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

// Proposal numbering of the reference's enum (cpu_rotormc.h:36-41).
enum RotorProposal { ROTOR_GIBBS_NE = 0, ROTOR_METROPOLIS_NE = 1, ROTOR_UNIFORM = 2, ROTOR_TF = 3 };

struct RotorArgs {
    int n;
    int R;            // reads in this launch
    int r0;           // first read index (seeds, states, statistics are indexed r0 + local)
    int spb;          // sweeps per schedule step
    int nf;           // schedule steps
    int interval;     // snapshot every `interval` steps (0: none)
    int snapshots;    // snapshots per read: 1 + (nf - 1) / interval
    int lanes_log2;   // lanes per read = 1 << lanes_log2
    int rpb;          // reads per block
    int rpb_log2;     // ceil(log2(rpb)); per-read arrays are indexed (v << rpb_log2) | slot
    int threads;      // threads per block
    int csr_in_lds;
    int wave_per_read;  // rotor_wave_kernel: one wave per read (sequential angle proposals, rows <= 64)
    int row_max;        // longest adjacency row
    int row_min;        // shortest adjacency row
    int64_t m;        // couplers (final energy)
    int64_t m2;       // adjacency entries
    const double* h;
    const double* tf;
    const int* rowptr;
    const int* col;
    const double* J;
    const uint8_t* serial;  // row repeats a neighbour or holds the variable itself
    const int* cs;
    const int* ce;
    const double* cw;
    const double* hp;       // nf
    const double* hd;       // nf
    const double* tables;   // cos[256] then |sin|[256]
    const uint64_t* seeds;
    uint8_t* states;        // R x n angles, in/out
    double* energies;
    uint8_t* stats;         // (R, snapshots, n) or null
    unsigned long long* counters;  // per read: accepted proposals (or null)
    // rotor_run_kernel (sequential uniform proposals, rows <= 16): runs of
    // consecutive pairwise non-adjacent variables, first | length << 16
    int run_kernel;
    int nruns;
    const uint32_t* runs;
    // per run, 16 slots (member j < length, else member 0): row start, then
    // variable | degree << 16 | run length << 24
    const uint32_t* run_rows;
    const uint32_t* jump;   // the 64-step xorshift128+ jump table (kJumpEntries uint4)
    // LDS byte offsets
    unsigned off_rowinfo, off_col, off_J, off_dE, off_ang, off_runs, off_jump, lds_bytes;
};

// Longest run of rotor_run_kernel: its lanes are (member j, earlier draws o)
// pairs, o <= j, so a run of K members takes K (K + 1) / 2 lanes.
constexpr int kRotorRunMax = 10;

hipError_t launch_rotor(const RotorArgs& a, int randomize, int proposal, int nblocks, hipStream_t stream);

}  // namespace misa
