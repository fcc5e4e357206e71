// Steepest descent (greedy local search) for gfx950: one wave per read.
//
// MI355X counterpart of the reference's dense steepest-descent solver
// (dwave/samplers/greedy/src/descent.cpp:117-187): flip energies
// dE_v = (-2 s_v)(h_v + sum_u s_u J_vu) for the read's state, then repeatedly
// flip the variable with the most negative dE (lowest index on ties), negate
// its dE and subtract 4 s_best J s_u from each neighbour's dE, until no dE is
// negative.  Self-loops are constant energy and are skipped.
//
// Per wave: dE (float64) and the spins live in LDS; the minimum (value, index)
// of every 64-variable chunk lives in registers, chunk c on lane c when there
// are at most 64 chunks (n <= 4096), else in LDS with each lane holding the
// minimum of its chunks c = lane (mod 64).  A step is
//   1. argmin: DPP row reduction of the lane minima plus four readlanes; the
//      lowest lane holding the minimum (ballot) names the variable, since lane
//      order is index order;
//   2. flip and neighbour updates, one lane per entry of the padded row (rows
//      that name a neighbour twice are applied in order by one lane);
//   3. recompute the chunks the step touched, four per pass so that their
//      reductions overlap.
// All reductions run with the whole wave active (DPP reads neighbouring lanes).
// The final energy is the reference's sequential float64 sum (same order),
// fed 64 terms at a time by the whole wave.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "descent_engine.h"

namespace misa {
namespace {

constexpr int kNoIndex = 0x7fffffff;
// DPP controls: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:4, row_ror:8.
constexpr int kQuadSwap1 = 0xB1, kQuadSwap2 = 0x4E, kRowRor4 = 0x124, kRowRor8 = 0x128;

__device__ __forceinline__ double readlane_f64(double x, int l) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Minimum over the 64 lanes; NaN lanes never win (fmin keeps the number).
__device__ __forceinline__ double wave_min_f64(double x) {
    x = __builtin_fmin(x, dpp_f64<kQuadSwap1>(x));
    x = __builtin_fmin(x, dpp_f64<kQuadSwap2>(x));
    x = __builtin_fmin(x, dpp_f64<kRowRor4>(x));
    x = __builtin_fmin(x, dpp_f64<kRowRor8>(x));
    return __builtin_fmin(__builtin_fmin(readlane_f64(x, 0), readlane_f64(x, 16)),
                          __builtin_fmin(readlane_f64(x, 32), readlane_f64(x, 48)));
}

__device__ __forceinline__ int wave_min_i32(int x) {
    x = min(x, __builtin_amdgcn_mov_dpp(x, kQuadSwap1, 0xF, 0xF, false));
    x = min(x, __builtin_amdgcn_mov_dpp(x, kQuadSwap2, 0xF, 0xF, false));
    x = min(x, __builtin_amdgcn_mov_dpp(x, kRowRor4, 0xF, 0xF, false));
    x = min(x, __builtin_amdgcn_mov_dpp(x, kRowRor8, 0xF, 0xF, false));
    return min(min(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 16)),
               min(__builtin_amdgcn_readlane(x, 32), __builtin_amdgcn_readlane(x, 48)));
}

// e + q[0] + q[1] + ... + q[cnt-1], added in that order.  For a full 64 the
// loads run ahead of the adds, which form one dependent chain.
__device__ __forceinline__ double add_in_order(double e, const double* q, int cnt) {
    if (cnt == 64) {
        double b[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) b[k] = q[k];
#pragma unroll
        for (int k = 0; k < 64; ++k) e += b[k];
        // scheduling groups (LDS reads 0x100, VALU 0x2): 16 double reads
        // first, then 16 adds per further 8 reads of two doubles
        __builtin_amdgcn_sched_group_barrier(0x100, 16, 0);
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
            __builtin_amdgcn_sched_group_barrier(0x2, 16, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x2, 32, 0);
    } else {
        for (int k = 0; k < cnt; ++k) e += q[k];
    }
    return e;
}

// Next distinct chunk among the lanes in `pend` (lanes whose entry names
// chunk `ch`), or -1; removes every lane naming it.
__device__ __forceinline__ int take_chunk(uint64_t& pend, int ch) {
    if (!pend) return -1;
    const int c = __builtin_amdgcn_readlane(ch, __builtin_ctzll(pend));
    pend &= __ballot(ch != c);
    return c;
}

struct WaveState {
    double* dE;
    // More than 64 chunks: the chunk minima.  Otherwise one update slot per
    // chunk (value, index, step that wrote it).
    double* cmin;
    int* cidx;
    uint32_t* cstep;
    int8_t* s;
    int n, nc, lane;
    bool lane_chunks;  // chunk c's minimum lives on lane c
    double lv;         // this lane's chunk minimum
    int li;
};

// Recompute the minima of up to K chunks (entries < 0 are unused).
template <int K>
__device__ __forceinline__ void update_chunks(WaveState& w, const int (&c)[K]) {
    double x[K], y[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int v = 64 * c[j] + w.lane;
        x[j] = (c[j] >= 0 && v < w.n) ? w.dE[v] : __builtin_inf();
        y[j] = x[j];
    }
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_fmin(y[j], dpp_f64<kQuadSwap1>(y[j]));
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_fmin(y[j], dpp_f64<kQuadSwap2>(y[j]));
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_fmin(y[j], dpp_f64<kRowRor4>(y[j]));
#pragma unroll
    for (int j = 0; j < K; ++j) y[j] = __builtin_fmin(y[j], dpp_f64<kRowRor8>(y[j]));
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (c[j] < 0) continue;
        const double m = __builtin_fmin(__builtin_fmin(readlane_f64(y[j], 0), readlane_f64(y[j], 16)),
                                        __builtin_fmin(readlane_f64(y[j], 32), readlane_f64(y[j], 48)));
        const uint64_t at = __ballot(x[j] == m);
        const int idx = at ? 64 * c[j] + __builtin_ctzll(at) : kNoIndex;
        if (w.lane_chunks) {
            if (w.lane == c[j]) {
                w.lv = m;
                w.li = idx;
            }
        } else if (w.lane == 0) {
            w.cmin[c[j]] = m;
            w.cidx[c[j]] = idx;
        }
    }
}

// Recompute the chunks whose bits are set (at most 64 chunks).
__device__ __forceinline__ void update_chunk_set(WaveState& w, uint64_t set) {
    if (!set) return;
    if (!(set & (set - 1))) {
        const int c[1] = {__builtin_ctzll(set)};
        update_chunks<1>(w, c);
        return;
    }
    while (set) {
        int c[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] = set ? __builtin_ctzll(set) : -1;
            set &= set - 1;
        }
        update_chunks<4>(w, c);
    }
}

// Lane minima over chunks c = lane (mod 64) from the LDS chunk minima.
__device__ __forceinline__ void rescan_chunks(WaveState& w) {
    __builtin_amdgcn_wave_barrier();
    w.lv = __builtin_inf();
    w.li = kNoIndex;
    for (int c = w.lane; c < w.nc; c += 64) {
        const double v = w.cmin[c];
        const int i = w.cidx[c];
        if (v < w.lv || (v == w.lv && i < w.li)) {
            w.lv = v;
            w.li = i;
        }
    }
}

// Row entry k of the chosen variable (k < re), or padding; CSR self entries
// become padding.
template <bool PADDED>
__device__ __forceinline__ void load_entry(const DescentArgs& a, int k, int re, int best, int& e, double& wk) {
    e = DESCENT_PAD;
    wk = 0.0;
    if (k < re) {
        if (PADDED) {
            e = a.ell_col[k];
            wk = a.ell_J[k];
        } else {
            const int u = a.col[k];
            e = u == best ? DESCENT_PAD : u;
            wk = a.J[k];
        }
    }
}

template <bool PADDED>
__global__ __launch_bounds__(256) void descent_kernel(DescentArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * a.waves + wave;
    if (r >= a.R) return;  // whole wave exits together (no block barrier below)
    const int n = a.n, nc = (n + 63) >> 6, D = a.ell_width;
    const size_t nde = (size_t)n;
    unsigned char* base = smem + (size_t)wave * a.wave_lds;
    WaveState w;
    w.dE = reinterpret_cast<double*>(base);
    w.cmin = reinterpret_cast<double*>(base + 8 * nde);
    w.cidx = reinterpret_cast<int*>(base + 8 * nde + 8 * (size_t)nc);
    w.cstep = reinterpret_cast<uint32_t*>(base + 8 * nde + 12 * (size_t)nc);
    w.s = reinterpret_cast<int8_t*>(base + 8 * nde + 16 * (size_t)nc);
    w.n = n;
    w.nc = nc;
    w.lane = lane;
    w.lane_chunks = nc <= 64;
    w.lv = __builtin_inf();
    w.li = kNoIndex;
    double* dE = w.dE;
    int8_t* s = w.s;
    int8_t* gs = a.states + (size_t)r * n;

    for (int v = lane; v < n; v += 64) s[v] = gs[v];
    for (int c = lane; c < nc; c += 64) w.cstep[c] = 0xFFFFFFFFu;  // no step writes this value
    __builtin_amdgcn_wave_barrier();
    // descent.cpp:41-62 and 133-137, one variable per lane
    for (int v = lane; v < n; v += 64) {
        double contrib = a.h[v];
        if (PADDED) {
            // padding lies inside the row, so every load is in bounds
            const int* rc = a.ell_col + (size_t)v * D;
            const double* rj = a.ell_J + (size_t)v * D;
            for (int k0 = 0; k0 < D; k0 += 8) {
                int e[8];
                double wv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    e[j] = k0 + j < D ? rc[k0 + j] : DESCENT_PAD;
                    wv[j] = k0 + j < D ? rj[k0 + j] : 0.0;
                }
                int8_t su[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) su[j] = s[e[j] >= 0 ? (e[j] & DESCENT_COL_MASK) : 0];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (e[j] >= 0) contrib += su[j] * wv[j];
            }
        } else {
            for (int k = a.rowptr[v]; k < a.rowptr[v + 1]; ++k) {
                const int u = a.col[k];
                if (u != v) contrib += s[u] * a.J[k];
            }
        }
        dE[v] = (-2 * s[v]) * contrib;
    }
    __builtin_amdgcn_wave_barrier();
    for (int c0 = 0; c0 < nc; c0 += 4) {
        const int c[4] = {c0, c0 + 1 < nc ? c0 + 1 : -1, c0 + 2 < nc ? c0 + 2 : -1, c0 + 3 < nc ? c0 + 3 : -1};
        update_chunks<4>(w, c);
    }
    if (!w.lane_chunks) rescan_chunks(w);

    uint32_t steps = 0;
    bool ready = false;  // the previous step already found the next minimum
    double mval = 0.0;
    int best = 0;
    while (steps < a.max_steps) {
        // descent.cpp:145-162: the first variable with the most negative dE
        if (!ready) {
            mval = wave_min_f64(w.lv);
            if (mval < 0.0) {
                if (w.lane_chunks)
                    best = __builtin_amdgcn_readlane(w.li, __builtin_ctzll(__ballot(w.lv == mval)));
                else
                    best = wave_min_i32(w.lv == mval ? w.li : kNoIndex);
            }
        }
        if (!(mval < 0.0)) break;
        ready = false;
        int rb, re;
        if (PADDED) {
            rb = best * D;
            re = rb + D;
        } else {
            rb = a.rowptr[best];
            re = a.rowptr[best + 1];
        }
        // The first pass's row entries are loaded before the flip so that the
        // loads overlap it.
        int e;
        double wk;
        load_entry<PADDED>(a, rb + lane, re, best, e, wk);
        const int sb = -s[best];  // the new spin of best
        __builtin_amdgcn_wave_barrier();
        // descent.cpp:166-171
        if (lane == 0) {
            s[best] = (int8_t)sb;
            dE[best] = -dE[best];
        }
        __builtin_amdgcn_wave_barrier();
        const int cb = best >> 6;
        const bool serial0 = PADDED ? __ballot(e >= 0 && (e & DESCENT_ROW_SERIAL)) != 0 : a.serial[best] != 0;
        if (w.lane_chunks && re - rb <= 16 && !serial0) {
            // One pass of at most 16 distinct neighbours (descent.cpp:173-181).
            // Chunks other than best's are updated from their entry: the
            // minimum moves to the entry when the entry now beats it (published
            // through the chunk's update slot), and the chunk is recomputed
            // when the entry was its minimum and grew, or when two entries
            // publish to it (the one whose write lost sees the other's index).
            const bool valid = e >= 0;
            const int u = e & DESCENT_COL_MASK;
            const int ch = valid ? (u >> 6) : -1;
            const double cv = __shfl(w.lv, ch & 63, 64);
            const int ci = __shfl(w.li, ch & 63, 64);
            double d0 = 0.0, d1 = 0.0;  // the entry's dE before and after
            if (valid) {
                const int su = s[u];
                d0 = dE[u];
                d1 = d0 - 4 * sb * wk * su;
                dE[u] = d1;
            }
            const bool cand = valid && ch != cb;
            bool again = cand && (cv != cv || (u == ci && !(d1 <= d0)));
            const bool moved = cand && !again && (u == ci || d1 < cv || (d1 == cv && u < ci));
            if (moved) {
                w.cmin[ch] = d1;
                w.cidx[ch] = u;
                w.cstep[ch] = steps;
            }
            __builtin_amdgcn_wave_barrier();
            const int vb = 64 * cb + lane;
            const double x = vb < n ? dE[vb] : __builtin_inf();
            if (moved && w.cidx[ch] != u) again = true;
            if (lane < nc && w.cstep[lane] == steps) {
                w.lv = w.cmin[lane];
                w.li = w.cidx[lane];
            }
            uint64_t redo_lanes = __ballot(again);
            if (redo_lanes) {
                uint64_t set = 1ull << cb;
                while (redo_lanes) {
                    const int c = __builtin_amdgcn_readlane(ch, __builtin_ctzll(redo_lanes));
                    set |= 1ull << c;
                    redo_lanes &= __ballot(ch != c);
                }
                update_chunk_set(w, set);
            } else {
                // Best's chunk and the next minimum together: two interleaved
                // reductions, over the chunk and over the other chunks' minima.
                const double y = lane == cb ? __builtin_inf() : w.lv;
                double rx = x, ry = y;
                rx = __builtin_fmin(rx, dpp_f64<kQuadSwap1>(rx));
                ry = __builtin_fmin(ry, dpp_f64<kQuadSwap1>(ry));
                rx = __builtin_fmin(rx, dpp_f64<kQuadSwap2>(rx));
                ry = __builtin_fmin(ry, dpp_f64<kQuadSwap2>(ry));
                rx = __builtin_fmin(rx, dpp_f64<kRowRor4>(rx));
                ry = __builtin_fmin(ry, dpp_f64<kRowRor4>(ry));
                rx = __builtin_fmin(rx, dpp_f64<kRowRor8>(rx));
                ry = __builtin_fmin(ry, dpp_f64<kRowRor8>(ry));
                rx = __builtin_fmin(__builtin_fmin(readlane_f64(rx, 0), readlane_f64(rx, 16)),
                                    __builtin_fmin(readlane_f64(rx, 32), readlane_f64(rx, 48)));
                ry = __builtin_fmin(__builtin_fmin(readlane_f64(ry, 0), readlane_f64(ry, 16)),
                                    __builtin_fmin(readlane_f64(ry, 32), readlane_f64(ry, 48)));
                const uint64_t ax = __ballot(x == rx);
                const uint64_t ay = __ballot(y == ry);
                const int ix = ax ? 64 * cb + __builtin_ctzll(ax) : kNoIndex;
                const int iy = ay ? __builtin_amdgcn_readlane(w.li, __builtin_ctzll(ay)) : kNoIndex;
                if (lane == cb) {
                    w.lv = rx;
                    w.li = ix;
                }
                mval = __builtin_fmin(rx, ry);
                best = rx < ry ? ix : (ry < rx ? iy : min(ix, iy));
                ready = true;
            }
            __builtin_amdgcn_wave_barrier();
        } else {
            bool first = true;
            // descent.cpp:173-181, 64 entries per pass; every chunk a pass
            // touches (and best's on the first) is recomputed.
            for (int k0 = rb; k0 < re; k0 += 64) {
                if (!first) load_entry<PADDED>(a, k0 + lane, re, best, e, wk);
                const bool serial = first ? serial0 : a.serial[best] != 0;
                if (!serial) {
                    if (e >= 0) {
                        const int u = e & DESCENT_COL_MASK;
                        dE[u] -= 4 * sb * wk * s[u];
                    }
                } else {
                    const int cnt = min(64, re - k0);
                    for (int j = 0; j < cnt; ++j) {
                        const int ej = __builtin_amdgcn_readlane(e, j);
                        const double wj = readlane_f64(wk, j);
                        if (ej >= 0 && lane == 0) {
                            const int u = ej & DESCENT_COL_MASK;
                            dE[u] -= 4 * sb * wj * s[u];
                        }
                    }
                }
                __builtin_amdgcn_wave_barrier();
                const int ch = e >= 0 ? ((e & DESCENT_COL_MASK) >> 6) : -1;
                uint64_t pend = __ballot(ch >= 0 && !(first && ch == cb));
                int c[4];
                c[0] = first ? cb : take_chunk(pend, ch);
                c[1] = take_chunk(pend, ch);
                c[2] = take_chunk(pend, ch);
                c[3] = take_chunk(pend, ch);
                update_chunks<4>(w, c);
                while (pend) {
                    c[0] = take_chunk(pend, ch);
                    c[1] = take_chunk(pend, ch);
                    c[2] = take_chunk(pend, ch);
                    c[3] = take_chunk(pend, ch);
                    update_chunks<4>(w, c);
                }
                first = false;
                __builtin_amdgcn_wave_barrier();
            }
            if (first) {  // a row without neighbours
                const int c[1] = {cb};
                update_chunks<1>(w, c);
            }
        }
        if (!w.lane_chunks) rescan_chunks(w);
        ++steps;
    }
    __builtin_amdgcn_wave_barrier();
    for (int v = lane; v < n; v += 64) gs[v] = s[v];

    if (lane == 0) a.steps[r] = steps;
}

// Final energies, descent.cpp:76-95: local fields in variable order, then
// couplers in input order, one float64 sum per read.  One wave per read with
// the read's spins in LDS: each 64 terms are formed by the whole wave and pass
// through LDS (two alternating halves) to lane 0, which adds them in order.
// Small LDS per wave keeps many reads' add chains in flight on each SIMD.
__global__ __launch_bounds__(256) void descent_energy_kernel(DescentArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 4 + wave;
    if (r >= a.R) return;
    const int n = a.n;
    unsigned char* base = smem + (size_t)wave * descent_energy_wave_lds(n);
    double* buf = reinterpret_cast<double*>(base);
    int8_t* s = reinterpret_cast<int8_t*>(base + 1024);
    const int8_t* gs = a.states + (size_t)r * n;
    for (int v = lane; v < n; v += 64) s[v] = gs[v];
    __builtin_amdgcn_wave_barrier();
    if (a.exact_sum) {
        // Every partial sum is exact, so the lanes add integer multiples of
        // 2^exact_exp in any order and the result equals the sequential sum.
        int64_t acc = 0;
        for (int v = lane; v < n; v += 64) acc += s[v] * a.h_q[v];
        for (int64_t c = lane; c < a.m; c += 64) acc += s[a.cs[c]] * s[a.ce[c]] * a.cw_q[c];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) a.energies[r] = ldexp((double)acc, a.exact_exp);
        return;
    }
    double energy = 0.0;
    int half = 0;
    double hv = lane < n ? a.h[lane] : 0.0;
    for (int v0 = 0; v0 < n; v0 += 64) {
        double* q = buf + 64 * half;
        const int v = v0 + lane;
        if (v < n) q[lane] = s[v] * hv;
        hv = v + 64 < n ? a.h[v + 64] : 0.0;
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) energy = add_in_order(energy, q, min(64, n - v0));
        __builtin_amdgcn_wave_barrier();
        half ^= 1;
    }
    const int64_t m = a.m;
    int iu = 0, iv = 0;
    double wc = 0.0;
    if (lane < m) {
        iu = a.cs[lane];
        iv = a.ce[lane];
        wc = a.cw[lane];
    }
    for (int64_t c0 = 0; c0 < m; c0 += 64) {
        double* q = buf + 64 * half;
        if (c0 + lane < m) q[lane] = s[iu] * wc * s[iv];
        const int64_t cn = c0 + 64 + lane;
        if (cn < m) {
            iu = a.cs[cn];
            iv = a.ce[cn];
            wc = a.cw[cn];
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) energy = add_in_order(energy, q, (int)min((int64_t)64, m - c0));
        __builtin_amdgcn_wave_barrier();
        half ^= 1;
    }
    if (lane == 0) a.energies[r] = energy;
}

}  // namespace

hipError_t launch_descent_steps(const DescentArgs& a, hipStream_t stream);

hipError_t launch_descent(const DescentArgs& a, hipStream_t stream) {
    if (a.R <= 0) return hipSuccess;
    hipError_t e = launch_descent_steps(a, stream);
    if (e != hipSuccess) return e;
    const int elds = 4 * (int)descent_energy_wave_lds(a.n);
    if (elds > 65536) {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(descent_energy_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, elds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(descent_energy_kernel, dim3((a.R + 3) / 4), dim3(256), elds, stream, a);
    return hipGetLastError();
}

hipError_t launch_descent_steps(const DescentArgs& a, hipStream_t stream) {
    const int lds = a.waves * a.wave_lds;
    const void* fn = a.ell_width > 0 ? reinterpret_cast<const void*>(descent_kernel<true>)
                                     : reinterpret_cast<const void*>(descent_kernel<false>);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    const int nblocks = (a.R + a.waves - 1) / a.waves;
    if (a.ell_width > 0)
        hipLaunchKernelGGL(descent_kernel<true>, dim3(nblocks), dim3(64 * a.waves), lds, stream, a);
    else
        hipLaunchKernelGGL(descent_kernel<false>, dim3(nblocks), dim3(64 * a.waves), lds, stream, a);
    return hipGetLastError();
}

}  // namespace misa
