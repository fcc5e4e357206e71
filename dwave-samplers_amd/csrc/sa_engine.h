// Internal interface between the C-ABI host code (api.cpp) and the HIP sweep
// kernels (sa_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace misa {

// Arguments of one simulated-annealing launch.  All pointers are device
// pointers.  Reads [r0, r0 + R) of the caller's arrays are processed; block b
// owns reads r0 + b*rpb ... r0 + b*rpb + rpb - 1.
struct SaArgs {
    int n;                 // variables
    int R;                 // reads in this launch
    int r0;                // first read index
    int lanes_log2;        // lanes cooperating on one read (power of two)
    int rpb;               // reads per block = 64 >> lanes_log2 (one wave per block)
    // Adjacency in coupler insertion order (CSR).
    const int* rowptr;     // [n+1]
    const int* col;        // [2m]
    const double* Jd;      // [2m] weights (real mode)
    const int16_t* Ji;     // [2m] weights (integer modes)
    const uint8_t* serial_row;  // [n] 1 if the row repeats a neighbour or holds v itself
    const double* h;       // [n]
    // Couplers in input order (final energy, cpu_sa.cpp:210-227).
    int64_t m;
    const int* cs;
    const int* ce;
    const double* cw;
    // Schedule.
    int nbetas;
    int sweeps_per_beta;
    const double* betas;   // [nbetas]
    const double* thr;     // [nbetas] 44.36142 / beta (real mode)
    const int* kthr;       // [nbetas] integer modes: skip iff dE/2 >= kthr
    const uint64_t* ttab;  // [nbetas][2*fmax+1] integer modes: accept iff rand < T
    int fmax;
    // Reads.
    int8_t* states;        // [R_total][n] in/out
    double* energies;      // [R_total]
    const uint64_t* seeds; // [R_total]
    unsigned long long* counters;  // [R_total][2] or null
    void* gstate;          // per-read state in device memory (when not in LDS)
    // Integer LDS kernel only.
    int64_t m2;            // CSR entries (2m)
    const int* jnext;      // [n] sum of J over row-v entries whose column is (v+1) mod n
    const double* jnext_d;     // [n] J(v, (v+1) mod n) when one row-v entry holds it, else 0
    const uint8_t* next_adj;   // [n] 1 when exactly one row-v entry has column (v+1) mod n
    int twidth;            // 2*fmax+1
    int waves_per_block;
    int rpb_log2;          // log2(rpb); rpb is a power of two in the integer kernel
    const uint32_t* rowinfo;  // [n+1] (rowptr[v] << 14) | (j1[v] & 0x7f) << 7 | (j2[v] & 0x7f)
    int pipelined;         // 1: sequential order runs the two-ahead pipelined loop (int8, n >= 3)
    unsigned lds_off_J_spins; // real kernel: byte offset of the spin array (fields at 0)
    unsigned lds_off_rowinfo; // byte offsets of the LDS regions (state words at 0)
    unsigned lds_off_col;
    unsigned lds_off_J;
    unsigned lds_off_T;
    unsigned lds_bytes;
    // Per-beta decision table indexed by the stored word w (integer kernel,
    // sequential order): lut[b][w - lut_lo] = 0 (skip, no draw), ~0 (flip
    // without a draw) or T (draw; flip iff rand < T).
    const uint64_t* lut;
    int lut_lo;
    int lut_width;
    int max_row;           // longest CSR row
    int reg_deg;           // every CSR row has this many entries, or -1
    // Segment kernel (float64, sequential order): the variables 0..n-1 cut into
    // nseg runs of consecutive variables with no coupler inside a run;
    // seg[0..nseg] are the run starts (seg[nseg] = n), bit 31 set on a run of
    // one variable with a serial row.  rowptr_pad: rowptr followed by 64
    // copies of rowptr[n] (wave-uniform reads past a run's end stay in
    // bounds).  Per-read state is read-major: field of v at
    // F[slot * fstride + v], spin at S[slot * sstride + v].
    const int* seg;
    const int* rowptr_pad;
    int nseg;
    // Run tiles (one read per wave, null otherwise): run r's members' rows
    // padded to 1 << tile_plog2 entries, 64 / (1 << tile_plog2) members per
    // slot of 64 entries, from entry tile_off[4 r] (tile_off[4 r + 1..2]: row
    // bounds of the run's first member); padding: column -1, weight 0.
    const int* tile_off;
    const int* tile_col;
    const double* tile_J;
    int tile_plog2;
    int fstride;           // doubles per read's field row
    int sstride;           // bytes per read's spin row
    // Chain-run kernel (int8, sequential order): per variable 8 dwords, the
    // first (deg << 16) | (uint16_t)(-4 J1) with J1 the summed coupling of v
    // to v+1 (0 for v = n-1), then ((uint8_t)(-J) << 24) | (u << 4) for each
    // CSR entry (u, J) of its row (unused ones v << 4); runs of at most
    // kChainK variables, starts chain_runs[0..nchain_runs] (the last is n).
    // chain_jump: kJumpEntries uint4 (the RNG jump table), or null for the
    // variant that steps the generator inside each run (the run kernel with
    // one read per wave: the kSegJumpSteps table staged at lds_off_jump, or
    // null for the wave's serial stream).
    // chain_ncrit: the most entries of one record that name a variable after
    // it and before the end of the next run (those entries come first).
    const uint32_t* chain_rec;
    // chain_slots: run-major slot records (rows of at most 6 entries, or
    // null): slot (run r, lane k) at 8 (16 r + k), dword 7 = 16 v of its
    // member v, bit 31 set for a lane past the run.
    const uint32_t* chain_slots;
    const int* chain_runs;
    int nchain_runs;
    int chain_ncrit;
    const uint32_t* chain_jump;
    // chain_lut: per beta and stored word (as lut), 4 dwords: -T low and
    // high, T high, and kChainDraw when the entry draws or kChainAlways when
    // it flips without a draw (T = ~0); 0 otherwise.
    const uint32_t* chain_lut;
    // Few-threshold form (null when some beta has more than kChainFewT
    // distinct thresholds T over the words a variable can hold): chain_thr
    // holds per beta kChainFewT thresholds (unused ones 0), chain_lut2 per
    // beta and stored word 2 dwords: the shift 16 a of threshold a's bits
    // (48 when the entry does not draw) and the class bits (kChainDraw,
    // kChainAlways or 0).
    const uint64_t* chain_thr;
    const uint32_t* chain_lut2;
    // Dense kernel (float64, sequential order, n <= 128): the couplings as n
    // rows of 64 dense_nr doubles, J(v, u) at row v, column u, zero where v
    // and u are not coupled.
    const double* dense;
    int dense_nr;
    unsigned lds_off_rng;   // per-read RNG scratch or output rings of the chain-run kernel
    unsigned lds_off_jump;  // jump table (ring variant)
};

// Kernel variant selectors.
enum Mode { MODE_INT8 = 1, MODE_INT16 = 2, MODE_REAL = 3 };

// Bytes of per-read state for n variables in the given mode.
__host__ __device__ inline size_t state_bytes_per_read(int mode, int n) {
    switch (mode) {
        case MODE_INT8: return (size_t)n;
        case MODE_INT16: return (size_t)n * 2;
        default: return (size_t)n * 9;  // f64 field + int8 spin
    }
}

// Launches the SA sweep kernel.  lds_state selects LDS-resident per-read
// state (dynamic LDS of rpb * state_bytes_per_read bytes) or device memory
// (args.gstate).  Returns a hipError_t.
hipError_t launch_sa(const SaArgs& args, int mode, int order, int proposal,
                     bool lds_state, int nblocks, hipStream_t stream);

// Integer-mode fast kernel: CSR staged in LDS (or read from device memory),
// packed biased fields in LDS updated with return-less LDS atomics, next
// word prefetched.  Geometry in args (waves_per_block, lanes_log2, rpb and
// the lds_off_* offsets).  mode is MODE_INT8 or MODE_INT16.
hipError_t launch_sa_int_lds(const SaArgs& args, int mode, int order, int proposal, bool csr_in_lds,
                             int nblocks, hipStream_t stream);

// Real-mode (float64) LDS kernel: fields and spins in LDS, CSR staged in LDS
// (or read from device memory), return-less float64 LDS atomics for rows
// without repeated entries.  Geometry and LDS offsets in args.
hipError_t launch_sa_real_lds(const SaArgs& args, int order, int proposal, bool csr_in_lds, int nblocks,
                              hipStream_t stream);

// Segment kernel for float64 problems in sequential order: each read's lanes
// decide a whole run of mutually non-adjacent variables at once (lanes per
// read 8..64 as 1 << lanes_log2; at most seg_kmax(lanes_log2) variables per
// run).  Geometry and LDS offsets in args.
hipError_t launch_sa_seg(const SaArgs& args, int proposal, bool csr_in_lds, int nblocks, hipStream_t stream);

// Chain-run kernel for int8 problems in sequential order (16 lanes per read,
// 4 waves of 4 reads per block): each run of consecutive variables, where a
// variable is coupled to no earlier one of its run but its predecessor, is
// decided by one lane per variable and a serial pass over the run.  entries:
// neighbour entries per variable read from chain_rec (4, 6 or 7).  LDS: 16 n
// bytes of words, then the per-wave tables at lds_off_T and the RNG scratch
// at lds_off_rng.
constexpr int kChainK = 15;
// Class bits of a chain-run table entry: draw, or flip without a draw (every
// outcome bit set).
constexpr uint32_t kChainDraw = 1u << 15;
constexpr uint32_t kChainAlways = (1u << kChainK) - 1u;
constexpr int kChainRecDwords = 8;
// Most distinct draw thresholds per beta in the few-threshold form.
constexpr int kChainFewT = 3;
// RNG scratch entries per read: the state words n_{-1}, n_0 at a run's start
// and the kChainK after them, rounded up to an even count.
constexpr int kChainRngStride = ((kChainK + 3) / 2) * 2;
// Ring variant: each read's xorshift128+ outputs are made kRingBatch at a time
// by the 64 lanes of its wave (kRingM consecutive outputs per lane, after
// which a lane's sub-stream jumps kJumpSteps steps to its place in the next
// block) into a ring of kRingSize entries, the first 16 mirrored past its
// end (kRingStride entries per read).  The jump table has 32 nibbles x 16
// values of 128 bits.
constexpr int kRingM = 4;
constexpr int kRingBatch = 64 * kRingM;
constexpr int kRingSize = 2 * kRingBatch;
constexpr int kRingStride = kRingSize + 16;
constexpr int kJumpSteps = kRingBatch - kRingM;
constexpr int kJumpEntries = 32 * 16;
// Run kernel with one read per wave: lane j's random sub-stream yields the
// outputs j, j + 64, j + 128, ... of the read's stream (a 64-step jump).
constexpr int kSegJumpSteps = 64;
hipError_t launch_sa_chain(const SaArgs& args, int entries, int nblocks, hipStream_t stream);

// Dense kernel for float64 problems in sequential order with n <= 128 and no
// repeated couplers or self-loops: one wave per read (waves_per_block reads
// per block), fields in registers, the dense coupling matrix (args.dense,
// args.dense_nr) staged in lds_bytes of LDS.
hipError_t launch_sa_dense(const SaArgs& args, int proposal, int nblocks, hipStream_t stream);

// Longest run the segment kernel takes with 1 << lanes_log2 lanes per read.
__host__ __device__ constexpr int seg_kmax(int lanes_log2) {
    return lanes_log2 >= 6 ? 32 : (1 << lanes_log2) - 1;
}

// Work item of a batched launch: a block of reads of one problem, run by
// the segment body (kind 1) or the pipelined float64 body (kind 0).
struct BatchItem {
    SaArgs args;
    int block;
    int csr_in_lds;
    int kind;
};

// Threads per block of the batched launch.
constexpr int kBatchThreads = 512;

// Batched float64 launch: a persistent grid of `grid` blocks of kBatchThreads
// threads takes items in order from a device work queue; queue points at two
// zeroed uint32 counters (next item, finished blocks), which the launch
// leaves zeroed again.  Dynamic LDS lds_bytes (the largest item's need).
hipError_t launch_sa_batch(const BatchItem* d_items, int nitems, unsigned* queue, int grid, unsigned lds_bytes,
                           int order, int proposal, hipStream_t stream);
// Blocks of the batched kernel one CU holds with lds_bytes of dynamic LDS.
hipError_t batch_blocks_per_cu(int order, int proposal, unsigned lds_bytes, int* out);

// Largest dynamic LDS the sweep kernel may request on the current device.
size_t max_dynamic_lds();

// Where one problem's float64 values live on the device, and how its float32
// inputs (h at h_off, couplers in input order at j_off) map onto them:
// csr_cpl[k] is the coupler of CSR entry k, jn_cpl[v] the coupler held by
// jnext_d[v] (or -1: jnext_d[v] = 0).
struct ValueMap {
    int n;
    int64_t m;
    int64_t h_off, j_off;
    double* h;
    double* cw;
    double* Jd;
    double* jnext_d;
    const int* csr_cpl;
    const int* jn_cpl;
};

// Writes every problem's values from h32 / J32; max_len bounds n and 2m.
hipError_t launch_values(const ValueMap* d_maps, int nmaps, int64_t max_len, const float* h32, const float* J32,
                         hipStream_t stream);
// out[r] = splitmix64 output r of seed, r < count.
hipError_t launch_read_seeds(uint64_t seed, int64_t count, uint64_t* out, hipStream_t stream);
// out[i] = +1 when the top bit of splitmix64 output i of (seed ^ kSpinSeedMix)
// is set, else -1.
constexpr uint64_t kSpinSeedMix = 0xD1B54A32D192ED03ull;
hipError_t launch_random_spins(uint64_t seed, int64_t count, int8_t* out, hipStream_t stream);

}  // namespace misa
