// C-ABI host side of the MI355X sweep engine (include/misamplers.h).
//
// Replaces the reference entry general_simulated_annealing
// (dwave/samplers/sa/src/cpu_sa.h:63-78, cpu_sa.cpp:253-362): validates the
// couplers with the reference's messages, builds the adjacency in coupler
// insertion order, builds the schedule tables, moves data to the device,
// launches the sweep kernels and returns the completed read count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host.h"
#include "misamplers.h"
#include "sa_engine.h"

namespace misa_host {

thread_local std::string g_err;

int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 256;
    return cus;
}

uint64_t splitmix_at(uint64_t seed, int64_t r) {
    uint64_t z = seed + (uint64_t)(r + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace misa_host

using namespace misa_host;


struct misa_sa_plan {
    misa_problem* p = nullptr;
    int spb = 1, order = 0, prop = 1, mode = MISA_MODE_REAL, lanes_req = 0;
    std::vector<double> betas, thr;
    std::vector<int> kthr;
    std::vector<uint64_t> ttab;
    std::vector<uint64_t> lut;  // per beta, indexed by stored word - lut_lo
    int lut_lo = 0, lut_width = 0;
    int fmax = 0;
    double *d_betas = nullptr, *d_thr = nullptr;
    int* d_kthr = nullptr;
    uint64_t* d_ttab = nullptr;
    uint64_t* d_lut = nullptr;
    uint32_t* d_jump = nullptr;  // chain-run ring variant's RNG jump table (built on first use)
    uint32_t* d_chain_lut = nullptr;  // chain-run kernel's decision table (built on first use)
    uint64_t* d_chain_thr = nullptr;  // its few-threshold form, when every beta has few thresholds
    uint32_t* d_chain_lut2 = nullptr;
    void* d_gstate = nullptr;
    size_t gstate_bytes = 0;
    // last launch geometry
    int lanes_log2 = 0, rpb = 64;
    bool lds_state = true;
    bool int_kernel = false;
    bool real_kernel = false;
    bool seg_kernel = false;
    bool chain_kernel = false;
    bool dense_kernel = false;
    int chain_form = 0;  // misa_sa_stats::chain_form of the last chain-run launch
    bool csr_in_lds = false;
    int waves_per_block = 1;
    ~misa_sa_plan() {
        dev_free(d_betas); dev_free(d_thr); dev_free(d_kthr); dev_free(d_ttab); dev_free(d_lut);
        dev_free(d_jump); dev_free(d_chain_lut); dev_free(d_chain_thr); dev_free(d_chain_lut2);
        if (d_gstate) (void)hipFree(d_gstate);
    }
};

namespace {

// Smallest r in [0, 2^64) for which accept(r) is false; accept must be true on
// a prefix of [0, 2^64) (it is monotone in r because u64 -> double rounding
// and multiplication by a positive factor are monotone).
template <typename Pred>
uint64_t first_reject(Pred accept) {
    if (!accept(0ull)) return 0ull;
    uint64_t lo = 0ull, hi = ~0ull;  // accept(lo) true
    if (accept(hi)) return ~0ull;    // cannot happen for the reference's rules
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (accept(mid)) lo = mid; else hi = mid;
    }
    return hi;
}

// Integer-mode tables.  For dE = 2k (k in [-fmax, fmax]) and each beta:
//   skip iff dE >= 44.36142/beta   (cpu_sa.cpp:131,140)   -> k >= kthr[b]
//   Metropolis (k > 0): accept iff exp(-dE*beta)*2^64 > (double)rand  (:154)
//   Gibbs:              accept iff 2^64 > (double)rand*(1+exp(dE*beta)) (:163)
// Both become "accept iff rand < T[b][k]" with T the first rejected rand,
// computed with the host libm exp, the function the reference calls.
void build_int_tables(misa_sa_plan* pl) {
    const int F = pl->fmax, W = 2 * F + 1;
    const int nb = (int)pl->betas.size();
    pl->kthr.assign(nb, 0);
    pl->ttab.assign((size_t)nb * W, 0ull);
    const double two64 = 18446744073709551616.0;
    for (int b = 0; b < nb; ++b) {
        const double beta = pl->betas[b];
        const double thr = pl->thr[b];
        int kt = F + 1;
        for (int k = -F; k <= F + 1; ++k) {
            if ((double)(2 * k) >= thr) { kt = k; break; }
        }
        pl->kthr[b] = kt;
        for (int k = -F; k <= F && k < kt; ++k) {
            const double de = (double)(2 * k);
            uint64_t T = 0;
            if (pl->prop == MISA_METROPOLIS) {
                if (k <= 0) continue;  // accepted without a draw
                const double P = exp(-de * beta) * two64;
                T = first_reject([&](uint64_t r) { return P > (double)r; });
            } else {
                const double Q = 1 + exp(de * beta);
                T = first_reject([&](uint64_t r) { return two64 > (double)r * Q; });
            }
            pl->ttab[(size_t)b * W + (k + F)] = T;
        }
    }
    // Word-indexed table for the sequential integer kernel: the stored word
    // is BIAS + 2 f + b (b = 1 for spin -1), so dE/2 = b ? f : -f.  Entry 0 =
    // skip without a draw; ~0 = flip without a draw (Metropolis, dE <= 0); any
    // other value is T (a draw, flip iff rand < T).  T is never 0 or ~0 for a
    // drawn entry (0 < exp(-dE*beta)*2^64 for dE < 44.36142/beta; T <= 2^64-1024).
    const int bias = pl->mode == MISA_MODE_INT8 ? 128 : 32768;
    pl->lut_lo = bias - 2 * F;
    pl->lut_width = 4 * F + 2;
    pl->lut.assign((size_t)nb * pl->lut_width, 0ull);
    for (int b = 0; b < nb; ++b) {
        for (int w = pl->lut_lo; w < pl->lut_lo + pl->lut_width; ++w) {
            const int f = ((w - bias) >> 1), bit = (w - bias) & 1;
            const int k = bit ? f : -f;
            uint64_t e = 0ull;
            if (k < pl->kthr[b]) {
                if (pl->prop == MISA_METROPOLIS && k <= 0) e = ~0ull;
                else e = pl->ttab[(size_t)b * W + (k + F)];
            }
            pl->lut[(size_t)b * pl->lut_width + (w - pl->lut_lo)] = e;
        }
    }
}

struct Geometry {
    int lanes_log2;
    int rpb;
    bool lds;
    int nblocks;
};

// Lanes per read: enough lanes that the reads fill about one wave per SIMD
// (4 per CU), at most 64; then fewer reads per wave if the LDS cannot hold them.
Geometry choose_geometry(const misa_sa_plan* pl, int R) {
    const int n = pl->p->n;
    const int cus = device_cus();
    const int target_waves = cus * 4;
    int L = 1;
    if (pl->lanes_req > 0) {
        L = 1;
        while (L < pl->lanes_req && L < 64) L <<= 1;
    } else {
        while (L < 64 && ((int64_t)R * L) / 64 < target_waves) L <<= 1;
    }
    const size_t per = misa::state_bytes_per_read(pl->mode, n);
    const size_t lmax = misa::max_dynamic_lds();
    while (L < 64 && (size_t)(64 / L) * per > lmax) L <<= 1;
    Geometry g;
    g.lanes_log2 = 0;
    while ((1 << g.lanes_log2) < L) g.lanes_log2++;
    g.rpb = 64 / L;
    g.lds = (size_t)g.rpb * per <= lmax;
    // Layout override for tests: keep the per-read state in device memory.
    if (const char* f = std::getenv("MISA_STATE_IN_DEVICE_MEMORY"))
        if (f[0] == '1') g.lds = false;
    g.nblocks = (R + g.rpb - 1) / g.rpb;
    return g;
}

struct IntGeometry {
    bool ok = false;
    int lanes_log2 = 0, rpb = 0, rpb_log2 = 0, waves = 1, nblocks = 0;
    bool csr = false;
    unsigned off_rowinfo = 0, off_col = 0, off_J = 0, off_T = 0, bytes = 0;
};

unsigned align16(size_t x) { return (unsigned)((x + 15) & ~(size_t)15); }

// Geometry of the integer LDS kernel: lanes per read as for the generic
// kernel; up to 4 waves per block share one staged copy of the adjacency.
// The per-read words, the row info, the per-wave acceptance rows and (when it
// fits) the adjacency must fit the LDS.
IntGeometry choose_int_geometry(const misa_sa_plan* pl, int R) {
    IntGeometry g;
    const misa_problem* p = pl->p;
    if (p->rowinfo.empty() || p->n < 2) return g;
    if (const char* f = std::getenv("MISA_STATE_IN_DEVICE_MEMORY"))
        if (f[0] == '1') return g;
    const int n = p->n;
    const size_t wsz = pl->mode == MISA_MODE_INT8 ? 1 : 2;
    const int cus = device_cus();
    int L = 1;
    if (pl->lanes_req > 0) {
        while (L < pl->lanes_req && L < 64) L <<= 1;
    } else {
        while (L < 64 && ((int64_t)R * L) / 64 < cus * 4) L <<= 1;
    }
    const size_t lmax = misa::max_dynamic_lds();
    const size_t m2 = 2 * (size_t)p->m;
    const size_t twidth = std::max<size_t>(2 * (size_t)pl->fmax + 1, (size_t)pl->lut_width);
    bool allow_csr = n <= 65536;
    if (const char* f = std::getenv("MISA_CSR_IN_DEVICE_MEMORY"))
        if (f[0] == '1') allow_csr = false;
    // Waves per block (sharing one staged adjacency): 4 by default; up to 8
    // with MISA_INT_WAVES (two waves per SIMD; slower on the lattice, where one
    // wave's instruction issue is the limit).
    int max_waves = 4;
    if (const char* f = std::getenv("MISA_INT_WAVES")) max_waves = std::max(1, std::min(8, std::atoi(f)));
    for (int csr = allow_csr ? 1 : 0; csr >= 0 && !g.ok; --csr) {
        for (int waves = max_waves; waves >= 1 && !g.ok; waves >>= 1) {
            for (int Lt = L; Lt <= 64 && !g.ok; Lt <<= 1) {
                const int rpw = 64 / Lt;
                const int rpb = waves * rpw;
                if (waves > 1 && (int64_t)(waves - 1) * rpw >= R) continue;  // a wave with no read
                if (((size_t)rpb * wsz) % 4 != 0) continue;  // rows of words must be dword aligned
                const unsigned offR = align16((size_t)rpb * n * wsz);
                const unsigned offC = offR + align16((size_t)(n + 1) * 4);
                // CSR arrays carry 64 spare entries: lanes past a row's end read them
                const unsigned offJ = csr ? offC + align16(2 * (m2 + 64)) : offC;
                const unsigned offT = csr ? offJ + align16((m2 + 64) * wsz) : offC;
                const size_t bytes = (size_t)offT + (size_t)waves * twidth * 8;
                if (bytes > lmax) continue;
                g.ok = true;
                g.csr = csr != 0;
                g.waves = waves;
                g.lanes_log2 = 0;
                while ((1 << g.lanes_log2) < Lt) g.lanes_log2++;
                g.rpb = rpb;
                g.rpb_log2 = 0;
                while ((1 << g.rpb_log2) < rpb) g.rpb_log2++;
                g.off_rowinfo = offR;
                g.off_col = offC;
                g.off_J = offJ;
                g.off_T = offT;
                g.bytes = (unsigned)bytes;
                g.nblocks = (R + rpb - 1) / rpb;
            }
        }
    }
    return g;
}

// MISA_CHAIN: "0" never takes the chain-run kernel, "1" takes it for every
// problem it can run; unset, it is taken when it is expected to be faster.
int chain_env() {
    const char* f = std::getenv("MISA_CHAIN");
    return f && (f[0] == '0' || f[0] == '1') ? f[0] - '0' : -1;
}

// The chain-run kernel runs int8 problems in sequential order whose rows hold
// at most 7 entries, when the caller asked for no other lane count.
bool chain_wanted(const misa_sa_plan* pl) {
    if (pl->mode != MISA_MODE_INT8 || pl->order != MISA_SEQUENTIAL) return false;
    if (pl->lanes_req > 0 && pl->lanes_req != 16) return false;
    if (chain_env() == 0) return false;
    if (const char* f = std::getenv("MISA_STATE_IN_DEVICE_MEMORY"))
        if (f[0] == '1') return false;
    const misa_problem* p = pl->p;
    return p->n >= 2 && p->n <= 65535 && p->maxdeg <= misa::kChainRecDwords - 1;
}

struct ChainGeometry {
    bool ok = false;
    bool ring = false;
    unsigned off_T = 0, off_rng = 0, off_jump = 0, bytes = 0;
    int nblocks = 0;
};

// LDS of a chain-run block: 16 words per variable, one decision-table row per
// wave, then either the output rings of the 16 reads (kRingStride entries
// each) and the jump table, or (when those do not fit, or MISA_CHAIN_RING is
// "0") 16 RNG scratch rows of kChainRngStride state words.  A run of the
// stepping variant costs about as much as eight variables of the pipelined
// integer kernel whatever its length, so by default the kernel is taken when
// the runs average at least 9 variables.  When the grid has no more blocks
// than CUs, a block asks for more than half the LDS so that no CU holds two.
ChainGeometry chain_geometry(const misa_sa_plan* pl, int R, int nruns) {
    ChainGeometry g;
    const int n = pl->p->n;
    if (nruns <= 0 || (chain_env() != 1 && (int64_t)nruns * 9 > n)) return g;
    g.off_T = align16((size_t)16 * n);
    g.off_rng = g.off_T + align16((size_t)4 * pl->lut_width * 16);
    const size_t lmax = misa::max_dynamic_lds();
    g.off_jump = g.off_rng + 16 * (unsigned)misa::kRingStride * 8;
    size_t need = (size_t)g.off_jump + (size_t)misa::kJumpEntries * 16;
    g.ring = need <= lmax;
    if (const char* f = std::getenv("MISA_CHAIN_RING"))
        if (f[0] == '0') g.ring = false;
    if (!g.ring) {
        g.off_jump = 0;
        need = (size_t)g.off_rng + 16 * (size_t)misa::kChainRngStride * 8;
    }
    if (need > lmax) return g;
    g.nblocks = (R + 15) / 16;
    size_t bytes = need;
    if (g.nblocks <= device_cus()) bytes = std::max(bytes, std::min(lmax, lmax / 2 + 1024));
    g.bytes = (unsigned)bytes;
    g.ok = true;
    return g;
}

// An xorshift128+ jump table (the chain-run ring variant's with kJumpSteps
// steps, the run kernel's with kSegJumpSteps): entry 16 c + v (4 dwords: s0
// low, s0 high, s1 low, s1 high) is the image after `steps` xorshift128+
// steps of the state whose only set bits are nibble v at bits 4c..4c+3 of
// s0 | s1 << 64.  The state update is linear over GF(2), so the images of a
// state's 32 nibbles XOR to the image of the state.
std::vector<uint32_t> xs_jump_table(int steps = misa::kJumpSteps) {
    uint64_t c0[128], c1[128];
    for (int b = 0; b < 128; ++b) {
        uint64_t s0 = b < 64 ? 1ull << b : 0ull, s1 = b < 64 ? 0ull : 1ull << (b - 64);
        for (int i = 0; i < steps; ++i) {
            const uint64_t t = s0 ^ (s0 << 23);
            const uint64_t ns1 = t ^ s1 ^ (t >> 17) ^ (s1 >> 26);
            s0 = s1;
            s1 = ns1;
        }
        c0[b] = s0;
        c1[b] = s1;
    }
    std::vector<uint32_t> tab((size_t)misa::kJumpEntries * 4);
    for (int c = 0; c < 32; ++c)
        for (int v = 0; v < 16; ++v) {
            uint64_t r0 = 0, r1 = 0;
            for (int j = 0; j < 4; ++j)
                if ((v >> j) & 1) {
                    r0 ^= c0[4 * c + j];
                    r1 ^= c1[4 * c + j];
                }
            uint32_t* e = &tab[(size_t)(16 * c + v) * 4];
            e[0] = (uint32_t)r0;
            e[1] = (uint32_t)(r0 >> 32);
            e[2] = (uint32_t)r1;
            e[3] = (uint32_t)(r1 >> 32);
        }
    return tab;
}


// Device data of the chain-run kernel for this plan: the problem's records
// and runs, the plan's decision table in the kernel's form, and the jump
// table.  Made when the plan is created, so that runs allocate nothing.
int prepare_chain(misa_sa_plan* pl, int* nruns) {
    *nruns = 0;
    if (int rc = pl->p->get_chain(nruns)) return rc;
    if (*nruns == 0) return MISA_OK;
    if (!pl->d_jump)
        if (int rc = dev_upload(&pl->d_jump, xs_jump_table())) return rc;
    if (!pl->d_chain_lut) {
        std::vector<uint32_t> cl(pl->lut.size() * 4);
        for (size_t i = 0; i < pl->lut.size(); ++i) {
            const uint64_t T = pl->lut[i], nT = 0ull - T;
            cl[4 * i] = (uint32_t)nT;
            cl[4 * i + 1] = (uint32_t)(nT >> 32);
            cl[4 * i + 2] = (uint32_t)(T >> 32);
            cl[4 * i + 3] = T == ~0ull ? misa::kChainAlways : (T != 0ull ? misa::kChainDraw : 0u);
        }
        if (int rc = dev_upload(&pl->d_chain_lut, cl)) return rc;
        // Few-threshold form: per beta, the distinct T of the drawing words
        // that some variable can hold, when there are at most kChainFewT
        const int W = pl->lut_width, bias = 128;
        const size_t nb = pl->lut.size() / (size_t)W;
        std::vector<uint64_t> thr(nb * misa::kChainFewT, 0ull);
        std::vector<uint32_t> l2(pl->lut.size() * 2, 0u);
        bool few = !pl->p->chain_freach.empty();
        for (size_t b = 0; few && b < nb; ++b) {
            int D = 0;
            for (int i = 0; few && i < W; ++i) {
                const int f = (pl->lut_lo + i - bias) >> 1;
                const uint64_t T = pl->lut[b * W + i];
                uint32_t* e = &l2[2 * (b * W + i)];
                e[0] = 48u;
                if (f < -64 || f > 64 || !pl->p->chain_freach[f + 64] || T == 0ull) continue;
                if (T == ~0ull) {
                    e[1] = misa::kChainAlways;
                    continue;
                }
                int a = 0;
                while (a < D && thr[b * misa::kChainFewT + a] != T) ++a;
                if (a == D) {
                    if (D == misa::kChainFewT) few = false;
                    else thr[b * misa::kChainFewT + D++] = T;
                }
                e[0] = 16u * (uint32_t)a;
                e[1] = misa::kChainDraw;
            }
        }
        if (few) {
            if (int rc = dev_upload(&pl->d_chain_thr, thr)) return rc;
            if (int rc = dev_upload(&pl->d_chain_lut2, l2)) return rc;
        }
    }
    return MISA_OK;
}

// Geometry of the real-mode LDS kernel: float64 fields and int8 spins per
// read in LDS, row info, and (when it fits) the adjacency as u16 columns and
// float64 weights; up to 4 waves per block share the staged adjacency.
IntGeometry choose_real_geometry(const misa_sa_plan* pl, int R) {
    IntGeometry g;
    const misa_problem* p = pl->p;
    const int n = p->n;
    const int cus = device_cus();
    int L = 1;
    if (pl->lanes_req > 0) {
        while (L < pl->lanes_req && L < 64) L <<= 1;
    } else {
        while (L < 64 && ((int64_t)R * L) / 64 < cus * 4) L <<= 1;
    }
    if (const char* f = std::getenv("MISA_STATE_IN_DEVICE_MEMORY"))
        if (f[0] == '1') return g;
    const size_t lmax = misa::max_dynamic_lds();
    const size_t m2 = 2 * (size_t)p->m;
    bool allow_csr = n <= 65536;
    if (const char* f = std::getenv("MISA_CSR_IN_DEVICE_MEMORY"))
        if (f[0] == '1') allow_csr = false;
    for (int csr = allow_csr ? 1 : 0; csr >= 0 && !g.ok; --csr) {
        for (int waves = 4; waves >= 1 && !g.ok; waves >>= 1) {
            for (int Lt = L; Lt <= 64 && !g.ok; Lt <<= 1) {
                const int rpw = 64 / Lt;
                const int rpb = waves * rpw;
                if (waves > 1 && (int64_t)(waves - 1) * rpw >= R) continue;  // a wave with no read
                const unsigned offS = align16((size_t)rpb * n * 8);
                const unsigned offR = offS + align16((size_t)rpb * n);
                const unsigned offC = offR + align16((size_t)(n + 1) * 4);
                // staged adjacency: 64 spare entries past the last row
                const unsigned offJ = csr ? offC + align16(2 * (m2 + 64)) : offC;
                const size_t bytes = (size_t)offJ + (csr ? (m2 + 64) * 8 : 0);
                if (bytes > lmax) continue;
                g.ok = true;
                g.csr = csr != 0;
                g.waves = waves;
                g.lanes_log2 = 0;
                while ((1 << g.lanes_log2) < Lt) g.lanes_log2++;
                g.rpb = rpb;
                g.rpb_log2 = 0;
                while ((1 << g.rpb_log2) < rpb) g.rpb_log2++;
                g.off_rowinfo = offR;
                g.off_col = offC;
                g.off_J = offJ;
                g.off_T = offS;  // spins
                g.bytes = (unsigned)std::max<size_t>(bytes, 16);
                g.nblocks = (R + rpb - 1) / rpb;
            }
        }
    }
    return g;
}

// Geometry of the segment kernel (float64, sequential order): per read the
// fields (fstride doubles, odd so the reads of a wave fall on different
// banks) and spins (sstride bytes); shared by the block, when chosen, the
// adjacency (u16 columns + f64 weights + 64 spare entries).  Reads per wave
// 64 / L with L >= 8 and L >= the longest row.
//   policy bit 0: stage the adjacency in LDS even when that costs up to half
//                 the reads a block could otherwise hold (else only when it
//                 costs at most a tenth);
//   policy bit 1: aim at two waves per SIMD before packing more reads into a
//                 wave (else: one read per wave when the block holds at most
//                 max_waves reads, then one wave per SIMD);
//   policy bits 2-3: at most 8 >> bits reads per wave (64, 32, 16 or 8 lanes
//                 per read at least).
// reads_cap bounds the reads per block (a plain launch spreads few reads
// over the CUs).
struct SegGeo {
    bool ok = false;
    int lg = 6, waves = 1, rpb = 1, nblocks = 0, fst = 1, sst = 4;
    bool csr = false;
    bool jump = false;  // one read per wave with room for the 64-step jump table
    unsigned offS = 0, offC = 0, offJ = 0, offJmp = 0, bytes = 0;
};

SegGeo seg_geometry(const misa_problem* p, int R, int policy, size_t lmax, int max_waves, int reads_cap) {
    SegGeo g;
    const int n = p->n;
    if (n < 1 || n > 65535 || R < 1 || p->maxdeg > 64) return g;
    const size_t m2 = 2 * (size_t)p->m;
    const int fst = n | 1;
    int sst = (n + 3) & ~3;
    if (((sst >> 2) & 1) == 0) sst += 4;
    const size_t per = (size_t)fst * 8 + (size_t)sst;
    const size_t csr_bytes = align16(2 * (m2 + 64)) + (m2 + 64) * 8;
    // 3 x 16 bytes of alignment slack between the regions
    auto fit = [&](size_t shared) -> int {
        return lmax > shared + 48 ? (int)std::min<size_t>((lmax - shared - 48) / per, 1 << 20) : 0;
    };
    const int rf_no = std::min(fit(0), R);
    int rf_csr = std::min(fit(csr_bytes), R);
    if (const char* f = std::getenv("MISA_CSR_IN_DEVICE_MEMORY"))
        if (f[0] == '1') rf_csr = 0;
    if (rf_no < 1) return g;
    const bool csr = rf_csr >= 1 && ((policy & 1) ? 2 * rf_csr >= rf_no : 10 * rf_csr >= 9 * rf_no);
    int rf = std::min(csr ? rf_csr : rf_no, std::max(1, reads_cap));
    int rpw = 8 >> ((policy >> 2) & 3);
    if (!(policy & 2) && rf <= max_waves) {
        rpw = 1;  // few reads per block: one wave each, the SIMDs' latency spread over waves
    } else {
        const int tw = (policy & 2) ? 8 : 4;  // target waves per block (4 SIMDs)
        while (rpw > 1 && (rf / rpw < tw || 64 / rpw < p->maxdeg)) rpw >>= 1;
    }
    g.lg = 6;
    for (int x = rpw; x > 1; x >>= 1) g.lg--;
    g.waves = std::min(max_waves, (rf + rpw - 1) / rpw);
    g.rpb = std::min(rf, g.waves * rpw);
    g.csr = csr;
    g.fst = fst;
    g.sst = sst;
    g.offS = align16((size_t)g.rpb * fst * 8);
    g.offC = g.offS + align16((size_t)g.rpb * sst);
    g.offJ = csr ? g.offC + align16(2 * (m2 + 64)) : g.offC;
    size_t bytes = (size_t)g.offJ + (csr ? (m2 + 64) * 8 : 0);
    if (bytes > lmax) return g;
    if (g.lg == 6 && align16(bytes) + (size_t)misa::kJumpEntries * 16 <= lmax) {
        g.jump = true;
        g.offJmp = (unsigned)align16(bytes);
        bytes = g.offJmp + (size_t)misa::kJumpEntries * 16;
    }
    g.bytes = (unsigned)std::max<size_t>(bytes, 16);
    g.nblocks = (R + g.rpb - 1) / g.rpb;
    g.ok = true;
    return g;
}

// The run tiles of a launch with one read per wave (MISA_SEG_TILES=0 keeps
// the CSR loads), and with them the per-lane random streams when the jump
// table has room in LDS (MISA_SEG_JUMP=0 keeps the wave's serial stream).
void set_tiles(misa::SaArgs& a, const misa_problem::Runs& runs, const SegGeo& g) {
    a.tile_off = nullptr;
    a.tile_col = nullptr;
    a.tile_J = nullptr;
    a.tile_plog2 = 0;
    a.chain_jump = nullptr;
    a.lds_off_jump = 0;
    if (a.lanes_log2 != 6 || runs.tile_plog2 < 0) return;
    if (const char* f = std::getenv("MISA_SEG_TILES"))
        if (f[0] == '0') return;
    a.tile_off = runs.d_tile_off;
    a.tile_col = runs.d_tile_col;
    a.tile_J = runs.d_tile_J;
    a.tile_plog2 = runs.tile_plog2;
    if (const char* f = std::getenv("MISA_SEG_JUMP"))
        if (f[0] == '0') return;
    if (g.jump) {
        a.chain_jump = seg_jump_table();
        a.lds_off_jump = a.chain_jump ? g.offJmp : 0;
    }
}

// Whether the segment kernel takes this problem: runs long enough to pay
// (MISA_SEG=0 never, 2 always when eligible; default: mean run >= 6, where the
// 15-regular problems measured it to beat the pipelined float64 loop).
bool seg_wanted(const misa_problem* p, const misa_problem::Runs& runs) {
    int mode = 1;
    if (const char* f = std::getenv("MISA_SEG")) mode = std::atoi(f);
    if (mode == 0) return false;
    if (mode == 2) return true;
    return (double)p->n >= 6.0 * (double)runs.nseg;
}

// Whether the dense kernel takes this problem (n <= 128 with no repeated
// coupler or self-loop, and the caller asked for no other lane count or
// kernel: MISA_SEG=2, MISA_CSR_IN_DEVICE_MEMORY and MISA_STATE_IN_DEVICE_MEMORY
// select other kernels; MISA_DENSE=0 never, 1 whenever it can run; default:
// rows of at least 16 entries on average).
bool dense_wanted(const misa_sa_plan* pl) {
    const misa_problem* p = pl->p;
    int mode = -1;
    if (const char* f = std::getenv("MISA_DENSE")) mode = std::atoi(f);
    if (mode == 0 || p->n < 2 || p->n > 128) return false;
    if (pl->lanes_req > 0 && pl->lanes_req != 64) return false;
    if (std::getenv("MISA_STATE_IN_DEVICE_MEMORY") || std::getenv("MISA_CSR_IN_DEVICE_MEMORY")) return false;
    if (const char* f = std::getenv("MISA_SEG"))
        if (std::atoi(f) == 2) return false;
    return mode == 1 || 2.0 * (double)p->m >= 16.0 * (double)p->n;
}

int validate_enums(int varorder, int proposal) {
    if (varorder != MISA_SEQUENTIAL && varorder != MISA_RANDOM)
        return fail(MISA_ERR_ARG, "unknown variable order");
    if (proposal != MISA_GIBBS && proposal != MISA_METROPOLIS)
        return fail(MISA_ERR_ARG, "Unknown proposal_acceptance_criteria");
    return MISA_OK;
}

}  // namespace

// The run kernels' 64-step jump table on the current device (host.h).
const uint32_t* misa_host::seg_jump_table() {
    static std::mutex mu;
    static std::vector<uint32_t*> per_dev;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, nullptr);
    if (!per_dev[dev]) {
        uint32_t* d = nullptr;
        if (dev_upload(&d, xs_jump_table(misa::kSegJumpSteps)) != MISA_OK) return nullptr;
        per_dev[dev] = d;
    }
    return per_dev[dev];
}

extern "C" {

const char* misa_last_error(void) { return g_err.c_str(); }

const char* misa_version(void) { return "misamplers 0.1.0 (gfx950)"; }

int misa_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

void misa_read_seeds(uint64_t seed, int64_t num_reads, uint64_t* out) {
    for (int64_t r = 0; r < num_reads; ++r) out[r] = splitmix_at(seed, r);
}

int misa_read_seeds_device(uint64_t seed, int64_t num_reads, uint64_t* d_out, void* stream) {
    if (num_reads < 0) return fail(MISA_ERR_ARG, "negative size");
    if (num_reads > 0 && !d_out) return fail(MISA_ERR_ARG, "null buffer");
    MISA_HIP(misa::launch_read_seeds(seed, num_reads, d_out, reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

void misa_random_spins(uint64_t seed, int64_t count, int8_t* out) {
    for (int64_t i = 0; i < count; ++i) out[i] = (splitmix_at(seed ^ misa::kSpinSeedMix, i) >> 63) ? 1 : -1;
}

int misa_random_spins_device(uint64_t seed, int64_t count, int8_t* d_out, void* stream) {
    if (count < 0) return fail(MISA_ERR_ARG, "negative size");
    if (count > 0 && !d_out) return fail(MISA_ERR_ARG, "null buffer");
    MISA_HIP(misa::launch_random_spins(seed, count, d_out, reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

int misa_problem_create(int num_vars, const double* h, int64_t num_starts,
                        const int32_t* coupler_starts, int64_t num_ends,
                        const int32_t* coupler_ends, int64_t num_weights,
                        const double* coupler_weights, misa_problem** out) {
    if (!out) return fail(MISA_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (num_vars < 0) return fail(MISA_ERR_ARG, "num_vars must be non-negative");
    if (!(num_starts == num_ends && num_starts == num_weights))
        return fail(MISA_ERR_LENGTHS, "coupler vectors have mismatched lengths");
    if (num_starts < 0) return fail(MISA_ERR_ARG, "negative coupler count");
    if (2 * num_starts > (int64_t)INT32_MAX - 1) return fail(MISA_ERR_UNSUPPORTED, "too many couplers");
    if ((num_vars > 0 && !h) || (num_starts > 0 && (!coupler_starts || !coupler_ends || !coupler_weights)))
        return fail(MISA_ERR_ARG, "null array");
    const int n = num_vars;
    const int64_t m = num_starts;
    for (int64_t c = 0; c < m; ++c) {
        const int u = coupler_starts[c], v = coupler_ends[c];
        if (u < 0 || v < 0 || u >= n || v >= n)
            return fail(MISA_ERR_INDEX, "coupler indexes contain an invalid variable");
    }
    auto* p = new misa_problem();
    p->n = n;
    p->m = m;
    p->h.assign(h, h + n);
    p->cs.assign(coupler_starts, coupler_starts + m);
    p->ce.assign(coupler_ends, coupler_ends + m);
    p->cw.assign(coupler_weights, coupler_weights + m);
    // CSR in insertion order: coupler c appends v to row u, then u to row v.
    p->rowptr.assign(n + 1, 0);
    for (int64_t c = 0; c < m; ++c) {
        p->rowptr[p->cs[c] + 1]++;
        p->rowptr[p->ce[c] + 1]++;
    }
    for (int v = 0; v < n; ++v) p->rowptr[v + 1] += p->rowptr[v];
    std::vector<int> fill(p->rowptr.begin(), p->rowptr.end() - (n >= 0 ? 1 : 0));
    p->col.assign(2 * m, 0);
    p->Jd.assign(2 * m, 0.0);
    for (int64_t c = 0; c < m; ++c) {
        const int u = p->cs[c], v = p->ce[c];
        p->col[fill[u]] = v; p->Jd[fill[u]] = p->cw[c]; fill[u]++;
        p->col[fill[v]] = u; p->Jd[fill[v]] = p->cw[c]; fill[v]++;
    }
    // Per-row facts: degree, repeated neighbours / self entries, |field| bound.
    p->serial.assign(n, 0);
    bool integer = true;
    double fmax = 0.0;
    std::vector<int> mark(n, -1);
    for (int v = 0; v < n; ++v) {
        const int kb = p->rowptr[v], ke = p->rowptr[v + 1];
        p->maxdeg = std::max(p->maxdeg, ke - kb);
        double bound = std::fabs(p->h[v]);
        if (!(std::isfinite(p->h[v]) && p->h[v] == std::nearbyint(p->h[v]))) integer = false;
        for (int k = kb; k < ke; ++k) {
            const int u = p->col[k];
            if (u == v || mark[u] == v) p->serial[v] = 1;
            mark[u] = v;
            const double J = p->Jd[k];
            if (!(std::isfinite(J) && J == std::nearbyint(J))) integer = false;
            bound += std::fabs(J);
        }
        fmax = std::max(fmax, bound);
    }
    p->integer_valued = integer;
    p->fmax = fmax;
    p->reg_deg = n > 0 ? p->rowptr[1] - p->rowptr[0] : -1;
    for (int v = 1; v < n && p->reg_deg >= 0; ++v)
        if (p->rowptr[v + 1] - p->rowptr[v] != p->reg_deg) p->reg_deg = -1;
    // The coupling of v to (v+1) mod n when exactly one row-v entry holds it
    // (the float64 kernel patches its read-ahead field of v+1 with it).
    p->jnext_d.assign(n, 0.0);
    p->next_adj.assign(n, 0);
    for (int v = 0; v < n; ++v) {
        const int nx = (v + 1) % n;
        int cnt = 0;
        for (int k = p->rowptr[v]; k < p->rowptr[v + 1]; ++k)
            if (p->col[k] == nx) {
                ++cnt;
                p->jnext_d[v] = p->Jd[k];
            }
        if (cnt == 1) p->next_adj[v] = 1;
        else p->jnext_d[v] = 0.0;
    }
    if (integer && fmax <= 255.0) {
        p->Ji.resize(2 * m);
        for (int64_t k = 0; k < 2 * m; ++k) p->Ji[k] = (int16_t)p->Jd[k];
        // Row info for the integer kernel: (rowptr[v] << 14) | (j1 & 0x7f) << 7 |
        // (j2 & 0x7f), with j1 / j2 the summed couplings of v to (v+1) mod n and
        // (v+2) mod n (7-bit signed: used by the pipelined int8 loop only).
        std::vector<int> j2(n, 0);
        p->jnext.assign(n, 0);
        for (int v = 0; v < n; ++v) {
            const int nx1 = (v + 1) % n, nx2 = (v + 2) % n;
            for (int k = p->rowptr[v]; k < p->rowptr[v + 1]; ++k) {
                if (p->col[k] == nx1) p->jnext[v] += (int)p->Ji[k];
                if (p->col[k] == nx2) j2[v] += (int)p->Ji[k];
            }
        }
        if (2 * m + 64 < ((int64_t)1 << 18)) {
            p->rowinfo.assign(n + 1, 0u);
            for (int v = 0; v <= n; ++v) {
                const uint32_t a1 = v < n ? ((uint32_t)p->jnext[v] & 0x7fu) : 0u;
                const uint32_t a2 = v < n ? ((uint32_t)j2[v] & 0x7fu) : 0u;
                p->rowinfo[v] = ((uint32_t)p->rowptr[v] << 14) | (a1 << 7) | a2;
            }
        }
    }
    *out = p;
    return MISA_OK;
}

void misa_problem_destroy(misa_problem* p) { delete p; }

int misa_problem_num_vars(const misa_problem* p) { return p ? p->n : 0; }
int64_t misa_problem_num_couplers(const misa_problem* p) { return p ? p->m : 0; }
int misa_problem_max_degree(const misa_problem* p) { return p ? p->maxdeg : 0; }

int misa_problem_auto_mode(const misa_problem* p) {
    if (!p) return MISA_MODE_REAL;
    if (p->integer_valued && p->fmax <= 63.0) return MISA_MODE_INT8;
    if (p->integer_valued && p->fmax <= 255.0) return MISA_MODE_INT16;
    return MISA_MODE_REAL;
}

int misa_sa_plan_create(const misa_problem* pc, int sweeps_per_beta, const double* beta_schedule,
                        int num_betas, int varorder, int proposal, int mode, int lanes_per_read,
                        misa_sa_plan** out) {
    if (!out || !pc) return fail(MISA_ERR_ARG, "null pointer");
    *out = nullptr;
    int rc = validate_enums(varorder, proposal);
    if (rc) return rc;
    if (sweeps_per_beta < 0 || num_betas < 0 || (num_betas > 0 && !beta_schedule))
        return fail(MISA_ERR_ARG, "bad schedule");
    auto* p = const_cast<misa_problem*>(pc);
    const int automode = misa_problem_auto_mode(p);
    if (mode == MISA_MODE_AUTO) mode = automode;
    if (mode == MISA_MODE_INT8 && !(p->integer_valued && p->fmax <= 63.0))
        return fail(MISA_ERR_UNSUPPORTED, "int8 mode needs integer biases with |field| <= 63");
    if (mode == MISA_MODE_INT16 && !(p->integer_valued && p->fmax <= 255.0))
        return fail(MISA_ERR_UNSUPPORTED, "int16 mode needs integer biases with |field| <= 255");
    if (mode < MISA_MODE_INT8 || mode > MISA_MODE_REAL) return fail(MISA_ERR_ARG, "unknown mode");
    if ((rc = p->ensure_device())) return rc;
    auto* pl = new misa_sa_plan();
    pl->p = p;
    pl->spb = sweeps_per_beta;
    pl->order = varorder;
    pl->prop = proposal;
    pl->mode = mode;
    pl->lanes_req = lanes_per_read;
    pl->betas.assign(beta_schedule, beta_schedule + num_betas);
    pl->thr.resize(num_betas);
    for (int b = 0; b < num_betas; ++b) pl->thr[b] = 44.36142 / pl->betas[b];  // cpu_sa.cpp:131
    if (mode != MISA_MODE_REAL) {
        pl->fmax = (int)p->fmax;
        build_int_tables(pl);
    } else {
        pl->kthr.assign(std::max(num_betas, 1), 0);
        pl->ttab.assign(1, 0ull);
    }
    if (pl->lut.empty()) pl->lut.assign(1, 0ull);
    // the run kernel's cuts for every run-length cap it may take, made here so
    // that runs of the plan allocate nothing on the device
    if (mode == MISA_MODE_REAL && varorder == MISA_SEQUENTIAL && p->n >= 1 && p->n <= 65535) {
        if (!seg_jump_table()) {
            delete pl;
            return fail(MISA_ERR_DEVICE, "could not upload the run kernel's jump table");
        }
        misa_problem::Runs runs;
        for (int lg = 3; lg <= 6; ++lg)
            if ((rc = p->get_runs(misa::seg_kmax(lg), &runs))) {
                delete pl;
                return rc;
            }
    }
    // the dense kernel's matrix
    if (mode == MISA_MODE_REAL && varorder == MISA_SEQUENTIAL) {
        int nr = 0;
        if ((rc = p->get_dense(&nr))) {
            delete pl;
            return rc;
        }
    }
    // and the chain-run kernel's data
    if (mode == MISA_MODE_INT8 && varorder == MISA_SEQUENTIAL) {
        int nruns = 0;
        if ((rc = prepare_chain(pl, &nruns))) {
            delete pl;
            return rc;
        }
    }
    if ((rc = dev_upload(&pl->d_betas, pl->betas)) || (rc = dev_upload(&pl->d_thr, pl->thr)) ||
        (rc = dev_upload(&pl->d_kthr, pl->kthr)) || (rc = dev_upload(&pl->d_ttab, pl->ttab)) ||
        (rc = dev_upload(&pl->d_lut, pl->lut))) {
        delete pl;
        return rc;
    }
    *out = pl;
    return MISA_OK;
}

void misa_sa_plan_destroy(misa_sa_plan* plan) { delete plan; }

int misa_sa_plan_info(const misa_sa_plan* pl, misa_sa_stats* info) {
    if (!pl || !info) return fail(MISA_ERR_ARG, "null pointer");
    std::memset(info, 0, sizeof(*info));
    info->mode = pl->mode;
    info->lanes_per_read = 1 << pl->lanes_log2;
    info->reads_per_block = pl->rpb;
    info->state_in_lds = pl->lds_state ? 1 : 0;
    info->kernel = pl->dense_kernel ? 6 : pl->chain_kernel ? 5 : pl->int_kernel ? 1 : (pl->seg_kernel ? 4 : (pl->real_kernel ? 2 : 0));
    info->csr_in_lds = pl->csr_in_lds ? 1 : 0;
    info->waves_per_block = pl->waves_per_block;
    info->chain_form = pl->chain_kernel ? pl->chain_form : 0;
    return MISA_OK;
}

int misa_sa_run_device(misa_sa_plan* pl, int8_t* d_states, double* d_energies,
                       const uint64_t* d_seeds, int num_samples, uint64_t* d_counters,
                       void* stream) {
    return misa_host::sa_run_device_impl(pl, d_states, d_energies, d_seeds, num_samples, d_counters, stream,
                                         nullptr, 0);
}

}  // extern "C"

int misa_host::sa_run_device_impl(misa_sa_plan* pl, int8_t* d_states, double* d_energies, const uint64_t* d_seeds,
                                  int num_samples, uint64_t* d_counters, void* stream, void* ext_state,
                                  size_t ext_bytes) {
    if (!pl) return fail(MISA_ERR_ARG, "null plan");
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    const misa_problem* p = pl->p;
    if (num_samples == 0 || p->n == 0) return MISA_OK;
    if (!d_states || !d_energies || !d_seeds) return fail(MISA_ERR_ARG, "null device buffer");
    Geometry geo = choose_geometry(pl, num_samples);
    pl->lanes_log2 = geo.lanes_log2;
    pl->rpb = geo.rpb;
    pl->lds_state = geo.lds;
    misa::SaArgs a{};
    a.n = p->n;
    a.R = num_samples;
    a.r0 = 0;
    a.lanes_log2 = geo.lanes_log2;
    a.rpb = geo.rpb;
    a.rowptr = p->d_rowptr;
    a.col = p->d_col;
    a.Jd = p->d_Jd;
    a.Ji = p->d_Ji;
    a.serial_row = p->d_serial;
    a.jnext_d = p->d_jnext_d;
    a.next_adj = p->d_next_adj;
    a.max_row = p->maxdeg;
    a.reg_deg = p->reg_deg;
    a.h = p->d_h;
    a.m = p->m;
    a.cs = p->d_cs;
    a.ce = p->d_ce;
    a.cw = p->d_cw;
    a.nbetas = (int)pl->betas.size();
    a.sweeps_per_beta = pl->spb;
    a.betas = pl->d_betas;
    a.thr = pl->d_thr;
    a.kthr = pl->d_kthr;
    a.ttab = pl->d_ttab;
    a.fmax = pl->fmax;
    a.states = d_states;
    a.energies = d_energies;
    a.seeds = d_seeds;
    a.counters = reinterpret_cast<unsigned long long*>(d_counters);
    a.gstate = nullptr;
    pl->chain_kernel = false;
    pl->dense_kernel = false;
    // (an allocation-free call takes the kernel only with the data the plan made)
    if (chain_wanted(pl) && (!ext_state || (pl->d_jump && pl->d_chain_lut))) {
        int nruns = 0;
        if (int rc = prepare_chain(pl, &nruns)) return rc;
        ChainGeometry cg = chain_geometry(pl, num_samples, nruns);
        if (cg.ok) {
            a.lanes_log2 = 4;
            a.rpb = 16;
            a.rpb_log2 = 4;
            a.waves_per_block = 4;
            a.lut = pl->d_lut;
            a.lut_lo = pl->lut_lo;
            a.lut_width = pl->lut_width;
            a.chain_rec = p->d_chain_rec;
            // MISA_CHAIN_SLOTS=0 keeps the per-variable records and run bounds
            const char* slenv = std::getenv("MISA_CHAIN_SLOTS");
            a.chain_slots = (slenv && slenv[0] == '0') ? nullptr : p->d_chain_slots;
            a.chain_runs = p->d_chain_runs;
            a.nchain_runs = nruns;
            a.chain_ncrit = p->chain_ncrit;
            a.chain_jump = cg.ring ? pl->d_jump : nullptr;
            a.chain_lut = pl->d_chain_lut;
            // MISA_CHAIN_FEW=0 keeps the per-lane comparisons
            const char* fewenv = std::getenv("MISA_CHAIN_FEW");
            const bool few = cg.ring && pl->d_chain_thr && !(fewenv && fewenv[0] == '0');
            a.chain_thr = few ? pl->d_chain_thr : nullptr;
            a.chain_lut2 = few ? pl->d_chain_lut2 : nullptr;
            a.lds_off_T = cg.off_T;
            a.lds_off_rng = cg.off_rng;
            a.lds_off_jump = cg.off_jump;
            a.lds_bytes = cg.bytes;
            pl->lanes_log2 = 4;
            pl->rpb = 16;
            pl->lds_state = true;
            pl->int_kernel = true;
            pl->chain_kernel = true;
            pl->chain_form = few ? 3 : cg.ring ? 2 : 1;
            pl->csr_in_lds = false;
            pl->waves_per_block = 4;
            MISA_HIP(misa::launch_sa_chain(a, p->maxdeg, cg.nblocks, reinterpret_cast<hipStream_t>(stream)));
            return MISA_OK;
        }
    }
    if (pl->mode != MISA_MODE_REAL) {
        IntGeometry ig = choose_int_geometry(pl, num_samples);
        if (ig.ok) {
            a.lanes_log2 = ig.lanes_log2;
            a.rpb = ig.rpb;
            a.m2 = 2 * p->m;
            a.jnext = p->d_jnext;
            a.twidth = 2 * pl->fmax + 1;
            a.waves_per_block = ig.waves;
            a.rpb_log2 = ig.rpb_log2;
            a.rowinfo = p->d_rowinfo;
            a.lds_off_rowinfo = ig.off_rowinfo;
            a.lds_off_col = ig.off_col;
            a.lds_off_J = ig.off_J;
            a.lds_off_T = ig.off_T;
            a.lds_bytes = ig.bytes;
            a.pipelined = (pl->mode == MISA_MODE_INT8 && p->n >= 3) ? 1 : 0;
            a.lut = pl->d_lut;
            a.lut_lo = pl->lut_lo;
            a.lut_width = pl->lut_width;
            pl->lanes_log2 = ig.lanes_log2;
            pl->rpb = ig.rpb;
            pl->lds_state = true;
            pl->int_kernel = true;
            pl->csr_in_lds = ig.csr;
            pl->waves_per_block = ig.waves;
            MISA_HIP(misa::launch_sa_int_lds(a, pl->mode, pl->order, pl->prop, ig.csr, ig.nblocks,
                                             reinterpret_cast<hipStream_t>(stream)));
            return MISA_OK;
        }
    }
    if (pl->mode == MISA_MODE_REAL && pl->order == MISA_SEQUENTIAL && dense_wanted(pl)) {
        int nr = 0;
        if (int rc = pl->p->get_dense(&nr)) return rc;
        const size_t bytes = (size_t)p->n * 64 * nr * sizeof(double);
        if (nr > 0 && bytes <= misa::max_dynamic_lds()) {
            // one wave per read, the reads spread over every CU
            const int W = std::min(16, std::max(1, (num_samples + device_cus() - 1) / device_cus()));
            a.waves_per_block = W;
            a.dense = p->d_dense;
            a.dense_nr = nr;
            a.lds_bytes = (unsigned)bytes;
            pl->lanes_log2 = 6;
            pl->rpb = W;
            pl->lds_state = false;
            pl->int_kernel = false;
            pl->real_kernel = true;
            pl->seg_kernel = false;
            pl->dense_kernel = true;
            pl->csr_in_lds = true;
            pl->waves_per_block = W;
            MISA_HIP(misa::launch_sa_dense(a, pl->prop, (num_samples + W - 1) / W,
                                           reinterpret_cast<hipStream_t>(stream)));
            return MISA_OK;
        }
    }
    if (pl->mode == MISA_MODE_REAL && pl->order == MISA_SEQUENTIAL && p->n <= 65535 &&
        !std::getenv("MISA_STATE_IN_DEVICE_MEMORY")) {
        // reads spread over every CU before they are packed into blocks
        const int cap = std::max(1, (num_samples + device_cus() - 1) / device_cus());
        int pol = 0;
        if (const char* f = std::getenv("MISA_SEG_POLICY")) pol = std::atoi(f) & 15;
        SegGeo sg = seg_geometry(p, num_samples, pol, misa::max_dynamic_lds(), 8, cap);
        misa_problem::Runs runs;
        if (sg.ok && pl->p->get_runs(misa::seg_kmax(sg.lg), &runs) == MISA_OK && seg_wanted(p, runs)) {
            a.lanes_log2 = sg.lg;
            a.rpb = sg.rpb;
            a.m2 = 2 * p->m;
            a.waves_per_block = sg.waves;
            a.lds_off_J_spins = sg.offS;
            a.lds_off_col = sg.offC;
            a.lds_off_J = sg.offJ;
            a.lds_bytes = sg.bytes;
            a.fstride = sg.fst;
            a.sstride = sg.sst;
            a.seg = runs.d_starts;
            a.rowptr_pad = p->d_rowptr_pad;
            a.nseg = runs.nseg;
            set_tiles(a, runs, sg);
            pl->lanes_log2 = sg.lg;
            pl->rpb = sg.rpb;
            pl->lds_state = true;
            pl->int_kernel = false;
            pl->real_kernel = true;
            pl->seg_kernel = true;
            pl->csr_in_lds = sg.csr;
            pl->waves_per_block = sg.waves;
            MISA_HIP(misa::launch_sa_seg(a, pl->prop, sg.csr, sg.nblocks, reinterpret_cast<hipStream_t>(stream)));
            return MISA_OK;
        }
    }
    pl->seg_kernel = false;
    if (pl->mode == MISA_MODE_REAL) {
        IntGeometry rg = choose_real_geometry(pl, num_samples);
        if (rg.ok) {
            a.lanes_log2 = rg.lanes_log2;
            a.rpb = rg.rpb;
            a.rpb_log2 = rg.rpb_log2;
            a.m2 = 2 * p->m;
            a.waves_per_block = rg.waves;
            a.lds_off_J_spins = rg.off_T;
            a.lds_off_rowinfo = rg.off_rowinfo;
            a.lds_off_col = rg.off_col;
            a.lds_off_J = rg.off_J;
            a.lds_bytes = rg.bytes;
            pl->lanes_log2 = rg.lanes_log2;
            pl->rpb = rg.rpb;
            pl->lds_state = true;
            pl->int_kernel = false;
            pl->real_kernel = true;
            pl->csr_in_lds = rg.csr;
            pl->waves_per_block = rg.waves;
            MISA_HIP(misa::launch_sa_real_lds(a, pl->order, pl->prop, rg.csr, rg.nblocks,
                                              reinterpret_cast<hipStream_t>(stream)));
            return MISA_OK;
        }
    }
    pl->int_kernel = false;
    pl->real_kernel = false;
    pl->csr_in_lds = false;
    pl->waves_per_block = 1;
    if (!geo.lds && ext_state) {
        // the caller's scratch (allocation-free entry point)
        const size_t need = (size_t)geo.nblocks * geo.rpb * misa::state_bytes_per_read(pl->mode, p->n);
        if (need > ext_bytes)
            return fail(MISA_ERR_UNSUPPORTED, "per-read state of " + std::to_string(need) +
                                                  " bytes exceeds the allocation-free scratch of " +
                                                  std::to_string(ext_bytes) + " bytes");
        a.gstate = ext_state;
    } else if (!geo.lds) {
        const size_t need = (size_t)geo.nblocks * geo.rpb * misa::state_bytes_per_read(pl->mode, p->n);
        if (need > pl->gstate_bytes) {
            if (pl->d_gstate) MISA_HIP(hipFree(pl->d_gstate));
            pl->d_gstate = nullptr;
            pl->gstate_bytes = 0;
            MISA_HIP(hipMalloc(&pl->d_gstate, need));
            pl->gstate_bytes = need;
        }
        a.gstate = pl->d_gstate;
    }
    MISA_HIP(misa::launch_sa(a, pl->mode, pl->order, pl->prop, geo.lds, geo.nblocks,
                             reinterpret_cast<hipStream_t>(stream)));
    return MISA_OK;
}

extern "C" {

int misa_sa_sample(const misa_problem* p, int8_t* states, double* energies, int num_samples,
                   int sweeps_per_beta, const double* beta_schedule, int num_betas,
                   const uint64_t* read_seeds, int varorder, int proposal, int mode,
                   int lanes_per_read, misa_interrupt_cb interrupt, void* interrupt_ctx,
                   misa_sa_stats* stats) {
    if (!p) return fail(MISA_ERR_ARG, "null problem");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (num_samples < 0) return fail(MISA_ERR_ARG, "num_samples must be non-negative");
    int rc = validate_enums(varorder, proposal);
    if (rc) return rc;
    const int n = p->n;
    // simulated_annealing.pyx:130-136: nothing to sample.
    if ((int64_t)num_samples * n == 0) {
        for (int r = 0; r < num_samples; ++r) energies[r] = 0.0;
        return num_samples;
    }
    if (!states || !energies || !read_seeds) return fail(MISA_ERR_ARG, "null buffer");
    const size_t total = (size_t)num_samples * n;
    for (size_t i = 0; i < total; ++i)
        if (states[i] != 1 && states[i] != -1) return fail(MISA_ERR_ARG, "initial states must be +1/-1 spins");
    misa_sa_plan* pl = nullptr;
    if ((rc = misa_sa_plan_create(p, sweeps_per_beta, beta_schedule, num_betas, varorder, proposal, mode,
                                  lanes_per_read, &pl)))
        return rc;
    int8_t* d_states = nullptr;
    double* d_en = nullptr;
    uint64_t* d_seeds = nullptr;
    uint64_t* d_cnt = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int done = 0;
    float total_ms = 0.f;
    auto cleanup = [&]() {
        if (d_states) (void)hipFree(d_states);
        if (d_en) (void)hipFree(d_en);
        if (d_seeds) (void)hipFree(d_seeds);
        if (d_cnt) (void)hipFree(d_cnt);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        delete pl;
    };
#define MISA_TRY(expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            cleanup();                                                                     \
            return fail(MISA_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                  \
    } while (0)
    MISA_TRY(hipMalloc(&d_states, total));
    MISA_TRY(hipMalloc(&d_en, sizeof(double) * num_samples));
    MISA_TRY(hipMalloc(&d_seeds, sizeof(uint64_t) * num_samples));
    // per-read flip counters only when statistics are asked for (the kernels
    // skip the counting otherwise)
    if (stats) MISA_TRY(hipMalloc(&d_cnt, sizeof(uint64_t) * 2 * num_samples));
    MISA_TRY(hipMemcpy(d_states, states, total, hipMemcpyHostToDevice));
    MISA_TRY(hipMemcpy(d_seeds, read_seeds, sizeof(uint64_t) * num_samples, hipMemcpyHostToDevice));
    MISA_TRY(hipEventCreate(&e0));
    MISA_TRY(hipEventCreate(&e1));
    // Without an interrupt callback all reads go in one launch; with one, in
    // chunks of about one wave per SIMD, polling after every read in order.
    int chunk = num_samples;
    if (interrupt) {
        Geometry geo = choose_geometry(pl, num_samples);
        chunk = std::max(geo.rpb * device_cus() * 4, 1);
    }
    std::vector<uint64_t> cnt((size_t)2 * num_samples, 0);
    for (int r0 = 0; r0 < num_samples;) {
        const int R = std::min(chunk, num_samples - r0);
        MISA_TRY(hipEventRecord(e0, nullptr));
        rc = misa_sa_run_device(pl, d_states + (size_t)r0 * n, d_en + r0, d_seeds + r0, R,
                                d_cnt ? d_cnt + 2 * (size_t)r0 : nullptr, nullptr);
        if (rc) {
            std::string keep = g_err;
            cleanup();
            return fail(rc, keep);
        }
        MISA_TRY(hipEventRecord(e1, nullptr));
        MISA_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        MISA_TRY(hipEventElapsedTime(&ms, e0, e1));
        total_ms += ms;
        MISA_TRY(hipMemcpy(states + (size_t)r0 * n, d_states + (size_t)r0 * n, (size_t)R * n, hipMemcpyDeviceToHost));
        MISA_TRY(hipMemcpy(energies + r0, d_en + r0, sizeof(double) * R, hipMemcpyDeviceToHost));
        if (d_cnt)
            MISA_TRY(hipMemcpy(cnt.data() + 2 * (size_t)r0, d_cnt + 2 * (size_t)r0, sizeof(uint64_t) * 2 * R,
                               hipMemcpyDeviceToHost));
        bool stop = false;
        for (int r = r0; r < r0 + R; ++r) {
            done = r + 1;
            if (interrupt && interrupt(interrupt_ctx)) {
                stop = true;
                break;
            }
        }
        r0 += R;
        if (stop) break;
    }
    if (stats) {
        for (int r = 0; r < done; ++r) {
            stats->flips += cnt[2 * (size_t)r];
            stats->flip_degree += cnt[2 * (size_t)r + 1];
        }
        stats->kernel_ms = total_ms;
        stats->mode = pl->mode;
        stats->lanes_per_read = 1 << pl->lanes_log2;
        stats->reads_per_block = pl->rpb;
        stats->state_in_lds = pl->lds_state ? 1 : 0;
        stats->kernel = pl->dense_kernel ? 6 : pl->chain_kernel ? 5 : pl->int_kernel ? 1 : (pl->seg_kernel ? 4 : (pl->real_kernel ? 2 : 0));
        stats->csr_in_lds = pl->csr_in_lds ? 1 : 0;
        stats->waves_per_block = pl->waves_per_block;
        stats->chain_form = pl->chain_kernel ? pl->chain_form : 0;
    }
    cleanup();
#undef MISA_TRY
    return done;
}


// ---------------------------------------------------------------------------
// Batched launch
// ---------------------------------------------------------------------------
}  // extern "C"

struct misa_batch {
    std::vector<misa_problem*> probs;  // not owned
    std::vector<int> reads;
    std::vector<int64_t> state_off, read_off;
    int64_t total_reads = 0, total_state = 0;
    int order = 0, prop = 1, policy = 0;
    std::vector<misa_sa_plan*> plans;  // per problem: schedule tables (float64 mode)
    // items of the last buffer set (host copy kept for the asynchronous upload)
    std::vector<misa::BatchItem> items;
    misa::BatchItem* d_items = nullptr;
    size_t d_items_cap = 0;
    unsigned* d_queue = nullptr;  // work-queue counters {next item, finished blocks}, zero between launches
    unsigned lds_bytes = 0;
    int grid = 0;
    int seg_items = 0;
    const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
    int key_policy = -1;
    hipStream_t last_stream = nullptr;  // stream of the last launch (its items may still be read)
    bool launched = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    // device-side value updates (misa_batch_set_values_device), built on first use
    misa::ValueMap* d_vmaps = nullptr;
    int* d_vidx = nullptr;
    int64_t vmax_len = 0;
    ~misa_batch() {
        if (launched) (void)hipStreamSynchronize(last_stream);
        for (auto* pl : plans) delete pl;
        if (d_items) (void)hipFree(d_items);
        if (d_queue) (void)hipFree(d_queue);
        if (d_vmaps) (void)hipFree(d_vmaps);
        if (d_vidx) (void)hipFree(d_vidx);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace {

// Geometry policies of a batch (bits 0-1: both bodies, bits 2-3: run body).
constexpr int kBatchPolicies = 16;

struct BatchGeo {
    bool ok = false;
    int rpb = 0, lanes_log2 = 0, rpb_log2 = 0;
    bool csr = false;
    unsigned offS = 0, offR = 0, offC = 0, offJ = 0, bytes = 0;
};

constexpr int kBatchWaves = misa::kBatchThreads / 64;

// Pipelined float64 body in a kBatchThreads block: rpb (a power of two) reads
// of L = min(64, kBatchThreads / rpb) lanes each.  Policy bit 0: adjacency in
// LDS first; bit 1: half the reads per block.
BatchGeo batch_geometry(const misa_problem* p, int R, int policy, size_t lmax) {
    BatchGeo g;
    const int n = p->n;
    const size_t m2 = 2 * (size_t)p->m;
    const bool csr_first = (policy & 1) != 0;
    const bool halve = (policy & 2) != 0;
    int cap = misa::kBatchThreads;
    while (cap > 1 && cap / 2 >= R) cap >>= 1;  // no more slots than reads (rounded up)
    for (int pass = 0; pass < 2 && !g.ok; ++pass) {
        const bool want_csr = csr_first ? pass == 0 : false;
        for (int rpb = cap; rpb >= 1 && !g.ok; rpb >>= 1) {
            const unsigned offS = align16((size_t)rpb * n * 8);
            const unsigned offR = offS + align16((size_t)rpb * n);
            const unsigned offC = offR + align16((size_t)(n + 1) * 4);
            const size_t plain = offC;
            const unsigned offJ = offC + align16(2 * (m2 + 64));  // 64 spare entries
            const size_t with_csr = (size_t)offJ + (m2 + 64) * 8;
            bool csr = false;
            if (want_csr) {
                if (n > 65536 || with_csr > lmax) continue;
                csr = true;
            } else {
                if (plain > lmax) continue;
                csr = n <= 65536 && with_csr <= lmax;  // stage it anyway when it fits
            }
            int use = rpb;
            if (halve && use > 1) use >>= 1;
            g.ok = true;
            g.rpb = use;
            g.rpb_log2 = 0;
            while ((1 << g.rpb_log2) < use) g.rpb_log2++;
            const int L = std::min(64, misa::kBatchThreads / use);  // few reads: one wave per read, idle waves
            g.lanes_log2 = 0;
            while ((1 << g.lanes_log2) < L) g.lanes_log2++;
            g.offS = align16((size_t)use * n * 8);
            g.offR = g.offS + align16((size_t)use * n);
            g.offC = g.offR + align16((size_t)(n + 1) * 4);
            g.offJ = g.offC + align16(2 * (m2 + 64));
            g.csr = csr && ((size_t)g.offJ + (m2 + 64) * 8 <= lmax);
            g.bytes = g.csr ? (unsigned)((size_t)g.offJ + (m2 + 64) * 8) : g.offC;
            g.bytes = std::max<unsigned>(g.bytes, 16);
        }
    }
    return g;
}

// LDS a batched block may use: the device maximum less the queue slot the
// kernel keeps in static LDS.
size_t batch_lds_max() { return misa::max_dynamic_lds() - 256; }

int batch_build_items(misa_batch* b, int8_t* d_states, double* d_en, const uint64_t* d_seeds, uint64_t* d_cnt,
                      hipStream_t st) {
    const size_t lmax = batch_lds_max();
    // the previous launch may still read the old items: wait for it before
    // the table is rewritten
    if (b->launched) MISA_HIP(hipStreamSynchronize(b->last_stream));
    b->items.clear();
    b->lds_bytes = 16;
    b->seg_items = 0;
    struct Cost { double c; size_t idx; };
    std::vector<Cost> order;
    for (size_t i = 0; i < b->probs.size(); ++i) {
        misa_problem* p = b->probs[i];
        misa_sa_plan* pl = b->plans[i];
        const int R = b->reads[i];
        if (R == 0 || p->n == 0) continue;
        misa::SaArgs a{};
        a.n = p->n;
        a.R = R;
        a.r0 = 0;
        a.rowptr = p->d_rowptr;
        a.col = p->d_col;
        a.Jd = p->d_Jd;
        a.Ji = p->d_Ji;
        a.serial_row = p->d_serial;
        a.jnext_d = p->d_jnext_d;
        a.next_adj = p->d_next_adj;
        a.max_row = p->maxdeg;
        a.reg_deg = p->reg_deg;
        a.h = p->d_h;
        a.m = p->m;
        a.m2 = 2 * p->m;
        a.cs = p->d_cs;
        a.ce = p->d_ce;
        a.cw = p->d_cw;
        a.nbetas = (int)pl->betas.size();
        a.sweeps_per_beta = pl->spb;
        a.betas = pl->d_betas;
        a.thr = pl->d_thr;
        a.kthr = pl->d_kthr;
        a.ttab = pl->d_ttab;
        a.states = d_states + b->state_off[i];
        a.energies = d_en + b->read_off[i];
        a.seeds = d_seeds + b->read_off[i];
        a.counters = d_cnt ? reinterpret_cast<unsigned long long*>(d_cnt + 2 * b->read_off[i]) : nullptr;
        a.waves_per_block = kBatchWaves;
        int kind = 0, rpb = 0, csr = 0;
        // the segment body for sequential order when the runs pay
        if (b->order == MISA_SEQUENTIAL && p->n <= 65535) {
            SegGeo sg = seg_geometry(p, R, b->policy, lmax, kBatchWaves, R);
            misa_problem::Runs runs;
            if (sg.ok && p->get_runs(misa::seg_kmax(sg.lg), &runs) == MISA_OK && seg_wanted(p, runs)) {
                kind = 1;
                rpb = sg.rpb;
                csr = sg.csr ? 1 : 0;
                a.lanes_log2 = sg.lg;
                a.rpb = sg.rpb;
                a.lds_off_J_spins = sg.offS;
                a.lds_off_col = sg.offC;
                a.lds_off_J = sg.offJ;
                a.lds_bytes = sg.bytes;
                a.fstride = sg.fst;
                a.sstride = sg.sst;
                a.seg = runs.d_starts;
                a.rowptr_pad = p->d_rowptr_pad;
                a.nseg = runs.nseg;
                set_tiles(a, runs, sg);
                b->seg_items++;
            }
        }
        if (kind == 0) {
            BatchGeo g = batch_geometry(p, R, b->policy, lmax);
            if (!g.ok)
                return fail(MISA_ERR_UNSUPPORTED, "problem " + std::to_string(i) + " (n=" + std::to_string(p->n) +
                                                      ") is too large for the batched engine's LDS layout");
            rpb = g.rpb;
            csr = g.csr ? 1 : 0;
            a.lanes_log2 = g.lanes_log2;
            a.rpb = g.rpb;
            a.rpb_log2 = g.rpb_log2;
            a.lds_off_J_spins = g.offS;
            a.lds_off_rowinfo = g.offR;
            a.lds_off_col = g.offC;
            a.lds_off_J = g.offJ;
            a.lds_bytes = g.bytes;
        }
        b->lds_bytes = std::max(b->lds_bytes, a.lds_bytes);
        const int nblk = (R + rpb - 1) / rpb;
        const double chain = (double)p->n * (double)a.nbetas * (double)a.sweeps_per_beta;
        for (int k = 0; k < nblk; ++k) {
            misa::BatchItem it;
            it.args = a;
            it.block = k;
            it.csr_in_lds = csr;
            it.kind = kind;
            order.push_back({chain, b->items.size()});
            b->items.push_back(it);
        }
    }
    // longest chains first: the work queue hands them out first, so the tail
    // holds the short ones
    std::stable_sort(order.begin(), order.end(), [](const Cost& x, const Cost& y) { return x.c > y.c; });
    std::vector<misa::BatchItem> sorted;
    sorted.reserve(b->items.size());
    for (auto& o : order) sorted.push_back(b->items[o.idx]);
    b->items.swap(sorted);
    if (b->items.size() > b->d_items_cap) {
        if (b->d_items) MISA_HIP(hipFree(b->d_items));
        b->d_items = nullptr;
        b->d_items_cap = 0;
        MISA_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_items), sizeof(misa::BatchItem) * b->items.size()));
        b->d_items_cap = b->items.size();
    }
    if (!b->items.empty())
        MISA_HIP(hipMemcpyAsync(b->d_items, b->items.data(), sizeof(misa::BatchItem) * b->items.size(),
                                hipMemcpyHostToDevice, st));
    // resident blocks of one launch: every CU holds as many as its LDS and
    // wave slots allow, but no more blocks than items
    int per_cu = 1;
    MISA_HIP(misa::batch_blocks_per_cu(b->order, b->prop, b->lds_bytes, &per_cu));
    b->grid = (int)std::min<int64_t>((int64_t)device_cus() * std::max(per_cu, 1), (int64_t)b->items.size());
    b->key[0] = d_states;
    b->key[1] = d_en;
    b->key[2] = d_seeds;
    b->key[3] = d_cnt;
    b->key_policy = b->policy;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return MISA_OK;
}

}  // namespace

extern "C" {

int misa_batch_create(int num_problems, const misa_problem* const* problems, const int* num_reads,
                      const int* sweeps_per_beta, const double* const* beta_schedules, const int* num_betas,
                      int varorder, int proposal, misa_batch** out) {
    if (!out) return fail(MISA_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (num_problems < 0 || (num_problems > 0 && (!problems || !num_reads || !sweeps_per_beta ||
                                                  !beta_schedules || !num_betas)))
        return fail(MISA_ERR_ARG, "bad batch arrays");
    int rc = validate_enums(varorder, proposal);
    if (rc) return rc;
    auto* b = new misa_batch();
    b->order = varorder;
    b->prop = proposal;
    // the work-queue counters start at zero; every launch leaves them zero
    if (hipMalloc(reinterpret_cast<void**>(&b->d_queue), 2 * sizeof(unsigned)) != hipSuccess ||
        hipMemset(b->d_queue, 0, 2 * sizeof(unsigned)) != hipSuccess) {
        delete b;
        return fail(MISA_ERR_DEVICE, "work-queue allocation failed");
    }
    for (int i = 0; i < num_problems; ++i) {
        if (!problems[i] || num_reads[i] < 0) {
            delete b;
            return fail(MISA_ERR_ARG, "bad problem or read count in batch");
        }
        misa_sa_plan* pl = nullptr;
        if ((rc = misa_sa_plan_create(problems[i], sweeps_per_beta[i], beta_schedules[i], num_betas[i], varorder,
                                      proposal, MISA_MODE_REAL, 0, &pl))) {
            delete b;
            return rc;
        }
        b->probs.push_back(const_cast<misa_problem*>(problems[i]));
        b->plans.push_back(pl);
        b->reads.push_back(num_reads[i]);
        b->state_off.push_back(b->total_state);
        b->read_off.push_back(b->total_reads);
        b->total_state += (int64_t)num_reads[i] * problems[i]->n;
        b->total_reads += num_reads[i];
    }
    *out = b;
    return MISA_OK;
}

void misa_batch_destroy(misa_batch* batch) { delete batch; }

int64_t misa_batch_state_offset(const misa_batch* b, int i) {
    return (b && i >= 0 && i < (int)b->state_off.size()) ? b->state_off[i] : -1;
}
int64_t misa_batch_read_offset(const misa_batch* b, int i) {
    return (b && i >= 0 && i < (int)b->read_off.size()) ? b->read_off[i] : -1;
}
int64_t misa_batch_total_reads(const misa_batch* b) { return b ? b->total_reads : 0; }
int64_t misa_batch_total_state(const misa_batch* b) { return b ? b->total_state : 0; }

int misa_batch_set_values_device(misa_batch* b, const float* d_h, const float* d_J, void* stream) {
    if (!b) return fail(MISA_ERR_ARG, "null batch");
    int64_t nh = 0, nj = 0;
    for (auto* p : b->probs) {
        nh += p->n;
        nj += p->m;
    }
    if ((nh > 0 && !d_h) || (nj > 0 && !d_J)) return fail(MISA_ERR_ARG, "null buffer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!b->d_vmaps && !b->probs.empty()) {
        // CSR entry -> coupler (the fill order of misa_problem_create) and the
        // coupler behind each jnext_d entry
        std::vector<int> idx;
        std::vector<std::pair<int64_t, int64_t>> at;  // per problem: offsets of csr_cpl, jn_cpl in idx
        for (auto* p : b->probs) {
            int rc = p->ensure_device();
            if (rc) return rc;
            const int n = p->n;
            std::vector<int> fill(p->rowptr.begin(), p->rowptr.end());
            std::vector<int> kc(2 * (size_t)p->m, 0);
            for (int64_t c = 0; c < p->m; ++c) {
                kc[fill[p->cs[c]]++] = (int)c;
                kc[fill[p->ce[c]]++] = (int)c;
            }
            std::vector<int> jn(n, -1);
            for (int v = 0; v < n; ++v)
                if (p->next_adj[v])
                    for (int k = p->rowptr[v]; k < p->rowptr[v + 1]; ++k)
                        if (p->col[k] == (v + 1) % n) jn[v] = kc[k];
            at.emplace_back((int64_t)idx.size(), (int64_t)idx.size() + (int64_t)kc.size());
            idx.insert(idx.end(), kc.begin(), kc.end());
            idx.insert(idx.end(), jn.begin(), jn.end());
            b->vmax_len = std::max<int64_t>(b->vmax_len, std::max<int64_t>(n, 2 * p->m));
        }
        int rc = dev_upload(&b->d_vidx, idx);
        if (rc) return rc;
        std::vector<misa::ValueMap> maps;
        int64_t ho = 0, jo = 0;
        for (size_t i = 0; i < b->probs.size(); ++i) {
            misa_problem* p = b->probs[i];
            misa::ValueMap mp{};
            mp.n = p->n;
            mp.m = p->m;
            mp.h_off = ho;
            mp.j_off = jo;
            mp.h = p->d_h;
            mp.cw = p->d_cw;
            mp.Jd = p->d_Jd;
            mp.jnext_d = p->d_jnext_d;
            mp.csr_cpl = b->d_vidx + at[i].first;
            mp.jn_cpl = b->d_vidx + at[i].second;
            maps.push_back(mp);
            ho += p->n;
            jo += p->m;
        }
        if ((rc = dev_upload(&b->d_vmaps, maps))) {
            dev_free(b->d_vidx);
            return rc;
        }
    }
    // the values must not change under a launch still running on another stream
    if (b->launched && b->last_stream != st) MISA_HIP(hipStreamSynchronize(b->last_stream));
    for (auto* p : b->probs) {  // host copies are stale from here on
        p->integer_valued = false;
        p->exact_sum = false;
    }
    MISA_HIP(misa::launch_values(b->d_vmaps, (int)b->probs.size(), b->vmax_len, d_h, d_J, st));
    return MISA_OK;
}

int misa_batch_set_policy(misa_batch* b, int policy) {
    if (!b || policy < 0 || policy >= kBatchPolicies)
        return fail(MISA_ERR_ARG, "policy must be 0.." + std::to_string(kBatchPolicies - 1));
    b->policy = policy;
    return MISA_OK;
}

int misa_batch_run_device(misa_batch* b, int8_t* d_states, double* d_energies, const uint64_t* d_seeds,
                          uint64_t* d_counters, void* stream, int use_graph) {
    if (!b) return fail(MISA_ERR_ARG, "null batch");
    if (b->total_reads == 0) return MISA_OK;
    if (!d_states || !d_energies || !d_seeds) return fail(MISA_ERR_ARG, "null device buffer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (use_graph && !st) return fail(MISA_ERR_ARG, "graph replay needs a non-null stream");
    int rc;
    if (b->key[0] != d_states || b->key[1] != d_energies || b->key[2] != d_seeds || b->key[3] != d_counters ||
        b->key_policy != b->policy || b->items.empty()) {
        if ((rc = batch_build_items(b, d_states, d_energies, d_seeds, d_counters, st))) return rc;
    }
    if (!use_graph) {
        MISA_HIP(misa::launch_sa_batch(b->d_items, (int)b->items.size(), b->d_queue, b->grid, b->lds_bytes, b->order,
                                       b->prop, st));
        b->last_stream = st;
        b->launched = true;
        return MISA_OK;
    }
    if (!b->exec || b->graph_stream != st) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        MISA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t le = misa::launch_sa_batch(b->d_items, (int)b->items.size(), b->d_queue, b->grid, b->lds_bytes,
                                              b->order, b->prop, st);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(st, &g);
        if (le != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(MISA_ERR_DEVICE, std::string("batched launch during capture: ") + hipGetErrorString(le));
        }
        if (ce != hipSuccess) return fail(MISA_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        b->graph = g;
        MISA_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->graph_stream = st;
    }
    MISA_HIP(hipGraphLaunch(b->exec, st));
    b->last_stream = st;
    b->launched = true;
    return MISA_OK;
}

int misa_batch_autotune(misa_batch* b, int trial_sweeps) {
    if (!b) return fail(MISA_ERR_ARG, "null batch");
    if (trial_sweeps < 1) trial_sweeps = 1;
    if (b->total_reads == 0) return b->policy;
    // trial copies of the plans: the middle of each schedule, trial_sweeps betas
    std::vector<misa_sa_plan*> saved = b->plans;
    std::vector<misa_sa_plan*> trial;
    int rc = MISA_OK;
    for (size_t i = 0; i < b->probs.size() && !rc; ++i) {
        const auto& bs = saved[i]->betas;
        std::vector<double> tb;
        for (int k = 0; k < trial_sweeps && !bs.empty(); ++k) tb.push_back(bs[(bs.size() / 2 + k) % bs.size()]);
        misa_sa_plan* pl = nullptr;
        rc = misa_sa_plan_create(b->probs[i], 1, tb.empty() ? nullptr : tb.data(), (int)tb.size(), b->order, b->prop,
                                 MISA_MODE_REAL, 0, &pl);
        trial.push_back(pl);
    }
    int8_t* d_st = nullptr;
    double* d_en = nullptr;
    uint64_t* d_sd = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = b->policy;
    float best_ms = 1e30f;
    const int keep_policy = b->policy;
    if (!rc) {
        std::vector<int8_t> hs((size_t)b->total_state);
        std::vector<uint64_t> hseed((size_t)b->total_reads);
        uint64_t x = 0x243F6A8885A308D3ull;
        for (auto& v : hs) { x ^= x << 13; x ^= x >> 7; x ^= x << 17; v = (x & 1) ? 1 : -1; }
        for (size_t i = 0; i < hseed.size(); ++i) hseed[i] = splitmix_at(12345, (int64_t)i);
        if (hipMalloc(&d_st, hs.size()) != hipSuccess || hipMalloc(&d_en, 8 * hseed.size()) != hipSuccess ||
            hipMalloc(&d_sd, 8 * hseed.size()) != hipSuccess ||
            hipMemcpy(d_sd, hseed.data(), 8 * hseed.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            rc = fail(MISA_ERR_DEVICE, "autotune allocation failed");
        } else {
            b->plans = trial;
            for (int pol = 0; pol < kBatchPolicies && !rc; ++pol) {
                b->policy = pol;
                b->items.clear();
                float ms = 0.f;
                for (int rep = 0; rep < 2 && !rc; ++rep) {  // first run warms up
                    if (hipMemcpy(d_st, hs.data(), hs.size(), hipMemcpyHostToDevice) != hipSuccess) {
                        rc = fail(MISA_ERR_DEVICE, "autotune copy failed");
                        break;
                    }
                    (void)hipEventRecord(e0, nullptr);
                    rc = misa_batch_run_device(b, d_st, d_en, d_sd, nullptr, nullptr, 0);
                    (void)hipEventRecord(e1, nullptr);
                    (void)hipEventSynchronize(e1);
                    (void)hipEventElapsedTime(&ms, e0, e1);
                }
                if (!rc && ms < best_ms) {
                    best_ms = ms;
                    best = pol;
                }
            }
            b->plans = saved;
        }
    }
    for (auto* pl : trial) delete pl;
    if (d_st) (void)hipFree(d_st);
    if (d_en) (void)hipFree(d_en);
    if (d_sd) (void)hipFree(d_sd);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    b->items.clear();
    b->key[0] = nullptr;
    b->policy = rc ? keep_policy : best;
    return rc ? rc : best;
}

int misa_batch_sample(misa_batch* b, int8_t* states, double* energies, const uint64_t* seeds, misa_sa_stats* stats) {
    if (!b) return fail(MISA_ERR_ARG, "null batch");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (b->total_reads == 0) return MISA_OK;
    if (!states || !energies || !seeds) return fail(MISA_ERR_ARG, "null buffer");
    for (int64_t i = 0; i < b->total_state; ++i)
        if (states[i] != 1 && states[i] != -1) return fail(MISA_ERR_ARG, "initial states must be +1/-1 spins");
    int8_t* d_st = nullptr;
    double* d_en = nullptr;
    uint64_t *d_sd = nullptr, *d_cnt = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MISA_OK;
    float ms = 0.f;
    std::vector<uint64_t> cnt((size_t)(2 * b->total_reads), 0);
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && !rc) rc = fail(MISA_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
        return rc == MISA_OK;
    };
    if (chk(hipMalloc(&d_st, (size_t)b->total_state), "hipMalloc") &&
        chk(hipMalloc(&d_en, 8 * (size_t)b->total_reads), "hipMalloc") &&
        chk(hipMalloc(&d_sd, 8 * (size_t)b->total_reads), "hipMalloc") &&
        chk(hipMalloc(&d_cnt, 16 * (size_t)b->total_reads), "hipMalloc") &&
        chk(hipMemcpy(d_st, states, (size_t)b->total_state, hipMemcpyHostToDevice), "hipMemcpy") &&
        chk(hipMemcpy(d_sd, seeds, 8 * (size_t)b->total_reads, hipMemcpyHostToDevice), "hipMemcpy") &&
        chk(hipEventCreate(&e0), "hipEventCreate") && chk(hipEventCreate(&e1), "hipEventCreate") &&
        chk(hipEventRecord(e0, nullptr), "hipEventRecord")) {
        rc = misa_batch_run_device(b, d_st, d_en, d_sd, d_cnt, nullptr, 0);
        if (!rc && chk(hipEventRecord(e1, nullptr), "hipEventRecord") &&
            chk(hipEventSynchronize(e1), "hipEventSynchronize") && chk(hipEventElapsedTime(&ms, e0, e1), "elapsed") &&
            chk(hipMemcpy(states, d_st, (size_t)b->total_state, hipMemcpyDeviceToHost), "hipMemcpy") &&
            chk(hipMemcpy(energies, d_en, 8 * (size_t)b->total_reads, hipMemcpyDeviceToHost), "hipMemcpy") &&
            chk(hipMemcpy(cnt.data(), d_cnt, 16 * (size_t)b->total_reads, hipMemcpyDeviceToHost), "hipMemcpy")) {
            if (stats) {
                for (int64_t r = 0; r < b->total_reads; ++r) {
                    stats->flips += cnt[2 * r];
                    stats->flip_degree += cnt[2 * r + 1];
                }
                stats->kernel_ms = ms;
                stats->mode = MISA_MODE_REAL;
                stats->kernel = 3;
                stats->state_in_lds = 1;
                stats->waves_per_block = kBatchWaves;
            }
        }
    }
    if (d_st) (void)hipFree(d_st);
    if (d_en) (void)hipFree(d_en);
    if (d_sd) (void)hipFree(d_sd);
    if (d_cnt) (void)hipFree(d_cnt);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    b->items.clear();
    b->key[0] = nullptr;
    return rc;
}

}  // extern "C"
