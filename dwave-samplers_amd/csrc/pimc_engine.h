// Discrete-imaginary-time path-integral engine: launch arguments shared by the
// host plan (pimc_api.cpp) and the HIP kernels (pimc_kernel.hip).
//
// Every read keeps its worldlines as one 32-bit word per site (bit k = spin of
// Trotter slice k, set = +1; P <= 32 slices).  A block holds 2^rpb_log2 reads,
// their worldlines interleaved per site (word (i << rpb_log2) | read) in LDS,
// next to one copy of the adjacency (self-loops removed, coupler order),
// the fields and the colour classes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

enum PimcKind { PIMC_TABLE = 0, PIMC_SIGNED = 1, PIMC_PLANES = 2, PIMC_CHAIN = 3 };

// Chain updates: at most this many qubits per chain and qubits x slices per
// chain (the segments of one chain update live in per-lane arrays).
constexpr int kChainQMax = 32;
constexpr int kChainNodes = 512;

enum PimcCut { PIMC_CUT_NONE = 0, PIMC_CUT_PROP = 1, PIMC_CUT_ALL = 2 };

struct PimcArgs {
    int n;          // sites
    int R;          // reads in this launch
    int P;          // Trotter slices, 1..32
    unsigned maskP; // low P bits
    int nsteps;     // schedule steps
    int spb;        // sweeps per step
    int interval;   // statistics every interval-th step (0: none)
    int ncol;       // statistics columns per read
    int rpb_log2;   // reads per block = 1 << rpb_log2
    int nclass;     // colour classes
    int m2;         // adjacency entries (2 per non-self coupler)
    int ntab;       // distinct coupling values (PIMC_TABLE)
    int jtab_in_lds;
    double J0;      // common |J| (PIMC_SIGNED)
    int Q;          // PIMC_CHAIN: qubits per chain (chains c < n / Q)
    double jchain;  // PIMC_CHAIN: coupling of consecutive qubits of a chain
    const double* h;
    const int* rowptr;            // n + 1
    const uint16_t* col;          // m2; PIMC_SIGNED: bit 15 set when J < 0
    const uint16_t* jidx;         // m2 indexes into jtab (PIMC_TABLE)
    const double* jtab;           // ntab
    const int* class_off;         // nclass + 1
    const uint16_t* class_sites;  // n, class by class (PIMC_CHAIN: chains, padded to n)
    const double* bs;             // per step: Hp / P
    const float* ginv;            // per step: ln 2 / ln(1 - tanh(Hd / P))
    const int* cut;               // per step: PimcCut
    const uint64_t* seeds;        // R
    uint32_t* W;                  // R x n worldlines, in/out
    int8_t* stats;                // R x ncol x n, or null
    uint64_t* flips;              // R: slice spins changed per read, or null
    // dynamic LDS layout (byte offsets)
    int off_h, off_jtab, off_W, off_rowptr, off_coff, off_col, off_jidx, off_sites;
    int lds_bytes;
    int threads;
};

// Anneal the launch's reads; one block per 2^rpb_log2 reads.
hipError_t launch_pimc(const PimcArgs& a, int kind, hipStream_t stream);

// energies[r] = E(slice 0 of read r) in the reference's summation order
// (localPIMC.cpp:56-73); states[r] = slice 0 as +-1 (either may be null).
hipError_t launch_pimc_project(const uint32_t* W, int R, int n, const double* h, int64_t m, const int* cs,
                               const int* ce, const double* cw, int8_t* states, double* energies,
                               hipStream_t stream);

}  // namespace misa
