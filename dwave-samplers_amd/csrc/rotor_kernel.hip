// Rotor-model annealing for gfx950.
//
// Same algorithm, float64 operation order and xorshift128+ stream per read as
// the reference core (dwave/samplers/sqa/src/cpu_rotormc.cpp:106-297): each
// variable is an angle i (theta = 2 pi i / 256) with cos / |sin| tables; a
// sweep proposes, per variable, either the reflection i + 128 (NonErgodic
// proposals, dE bookkeeping) or a new angle (uniform, or the transverse-field
// window) accepted on the change of the log measure
// -beta f (cos' - cos) + beta_tf t (|sin'| - |sin|), f the local field summed
// afresh in adjacency order, as the reference does.
//
// Block = reads_per_block reads, L lanes per read.  The lanes of a read run
// the same decisions (equal RNG states); on an accepted reflection they split
// the neighbour updates (rows that repeat a neighbour or hold the variable
// itself are applied in adjacency order by one lane).  LDS: the tables
// (4 KB), row info, the adjacency (u16 columns, f64 weights) when it fits,
// then per read dE (NonErgodic only) and the angles, indexed
// (v << log2 rpb) | slot.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rotor_engine.h"
#include "sa_engine.h"
#include "xorshift.h"

namespace misa {
namespace {

constexpr double kTwo64 = 18446744073709551616.0;  // (double)RANDMAX, cpu_rotormc.cpp:40

// exp(t) 2^64 > (double)x for t < 0 (an angle proposal's acceptance,
// cpu_rotormc.cpp:280-283), settled in single precision when a proven margin
// allows.  a = exp2f(t log2 e) 2^24: for |t| <= 87 the conversions cost at
// most 1.5e-5 in the exponent (relative error of a at most 1.1e-5, v_exp_f32
// about 1 ulp); with xh = x >> 40, x lies in [xh 2^40, (xh + 1) 2^40), so
// a (1 - 2^-14) > xh + 1 proves acceptance and a (1 + 2^-14) < xh rejection.
// Below t = -87 a is zero or subnormal while exp(t) 2^24 < 2^-101, so
// a (1 + 2^-14) < xh still proves rejection whenever xh >= 1.  Anything else
// (about 1e-4 of the draws) evaluates the double expression.
__device__ __forceinline__ bool accept_exp(const double t, const uint64_t x) {
    const float xh = (float)(uint32_t)(x >> 40);
    const float a = __builtin_amdgcn_exp2f((float)t * 1.44269504f) * 16777216.0f;
    if (__builtin_amdgcn_readfirstlane(a * 0.99993896f > xh + 1.0f ? 1 : 0)) return true;
    if (__builtin_amdgcn_readfirstlane(a * 1.00006104f < xh ? 1 : 0)) return false;
    return __builtin_amdgcn_readfirstlane(exp(t) * kTwo64 > (double)x ? 1 : 0) != 0;
}

// double -> int as the reference's x86-64 build computes int(x) (cvttsd2si):
// truncation, INT32_MIN for NaN and out-of-range values.
__device__ __forceinline__ int32_t trunc_i32(double x) {
    if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
    return (int32_t)x;
}

template <int ORDER, int PROP, bool CSR_LDS>
__global__ __launch_bounds__(256) void rotor_kernel(RotorArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool NE = PROP == ROTOR_GIBBS_NE || PROP == ROTOR_METROPOLIS_NE;
    const int lg = a.lanes_log2;
    const int L = 1 << lg;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int j = lane & (L - 1);
    const int slot = wave * (64 >> lg) + (lane >> lg);
    const int lrpb = a.rpb_log2;
    const int n = a.n;
    const int rlocal = blockIdx.x * a.rpb + slot;
    const bool own = slot < a.rpb;
    const bool act = own && rlocal < a.R;
    const int r = a.r0 + rlocal;
    double* ctab = reinterpret_cast<double*>(smem);
    double* stab = ctab + 256;
    uint32_t* RI = reinterpret_cast<uint32_t*>(smem + a.off_rowinfo);
    uint16_t* ccol = reinterpret_cast<uint16_t*>(smem + a.off_col);
    double* cJ = reinterpret_cast<double*>(smem + a.off_J);
    double* dE = reinterpret_cast<double*>(smem + a.off_dE);
    uint8_t* ang = smem + a.off_ang;

    for (int i = threadIdx.x; i < 512; i += blockDim.x) ctab[i] = a.tables[i];
    for (int v = threadIdx.x; v <= n; v += blockDim.x)
        RI[v] = (uint32_t)a.rowptr[v] | (v < n && a.serial[v] ? 0x80000000u : 0u);
    if constexpr (CSR_LDS)
        for (int64_t k = threadIdx.x; k < a.m2; k += blockDim.x) {
            ccol[k] = (uint16_t)a.col[k];
            cJ[k] = a.J[k];
        }
    if (own)
        for (int v = j; v < n; v += L) ang[(v << lrpb) | slot] = act ? a.states[(size_t)r * n + v] : (uint8_t)0;
    __syncthreads();

    auto col_at = [&](int k) -> int { return CSR_LDS ? (int)ccol[k] : a.col[k]; };
    auto J_at = [&](int k) -> double { return CSR_LDS ? cJ[k] : a.J[k]; };
    // get_marginal_state_fieldC, cpu_rotormc.cpp:58-79.  Entries are fetched
    // eight at a time (columns and weights, then the neighbours' angles, then
    // their cosines: three dependent LDS round trips per eight entries) and
    // added in adjacency order.
    // Branch-free batches: every load uses an index clamped to the row's last
    // entry (entry 0 when there are none; that buffer always exists), and
    // entries past the row select column 0 and are left out of the sum.
    auto field_from = [&](double f, int k0_start, int ke) -> double {
        const int last = ke > 0 ? ke - 1 : 0;
        for (int k0 = k0_start; k0 < ke; k0 += 8) {
            int c[8];
            double w[8], t[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = min(k0 + i, last);
                const int ck = col_at(k);
                const double wk = J_at(k);
                c[i] = k0 + i < ke ? ck : 0;
                w[i] = wk;
            }
            uint8_t an[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) an[i] = ang[(c[i] << lrpb) | slot];
#pragma unroll
            for (int i = 0; i < 8; ++i) t[i] = ctab[an[i]];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const double g = f + t[i] * w[i];
                f = k0 + i < ke ? g : f;
            }
        }
        return f;
    };
    auto field = [&](int v) -> double {
        return field_from(a.h[v], (int)(RI[v] & 0x7fffffffu), (int)(RI[v + 1] & 0x7fffffffu));
    };
    // Sequential order with the angle proposals: while one attempt decides, the
    // next variable's h, t, first eight row entries, their angles and cosines
    // are fetched; only the current variable's own angle can change in
    // between, and that entry is patched with the cosine of its new angle.
    constexpr bool PIPE = !NE && ORDER == 0;
    struct Ahead {
        int kb, ke;
        int c[8];
        double w[8], t[8];
        double h, tf;
    } nx;
    auto fetch_ahead = [&](int v) {
        nx.kb = (int)(RI[v] & 0x7fffffffu);
        nx.ke = (int)(RI[v + 1] & 0x7fffffffu);
        nx.h = a.h[v];
        nx.tf = a.tf[v];
        uint8_t an[8];
        const int last = nx.ke > 0 ? nx.ke - 1 : 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = min(nx.kb + i, last);
            const int ck = col_at(k);
            nx.w[i] = J_at(k);
            const bool in = nx.kb + i < nx.ke;
            nx.c[i] = in ? ck : -1;
            an[i] = ang[((in ? ck : 0) << lrpb) | slot];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) nx.t[i] = ctab[an[i]];
    };
    if (NE && own)
        for (int v = j; v < n; v += L) {
            const int e = (v << lrpb) | slot;
            dE[e] = field(v) * (-2 * ctab[ang[e]]);  // :136-144
        }
    __syncthreads();
    if (!act) return;

    uint64_t s0 = a.seeds[r];
    if (s0 == 0) s0 = ~0ull;  // :407-408
    uint64_t s1 = 0;
    for (int i = 0; i < 32; ++i) (void)xs_next(s0, s1);  // burn-in, :145-147
    unsigned long long accepted = 0;
    uint8_t* snap = (a.stats && a.interval > 0) ? a.stats + (size_t)r * a.snapshots * n : nullptr;
    int vprev = -1;          // the previous attempt's variable (PIPE)
    double cprev = 0.0;      // cos of its angle after that attempt
    if constexpr (PIPE) fetch_ahead(0);

    for (int bi = 0; bi < a.nf; ++bi) {
        const double beta = a.hp[bi], beta_tf = a.hd[bi];
        const double rat = beta_tf / beta;  // :157
        for (int sw = 0; sw < a.spb; ++sw) {
            for (int it = 0; it < n; ++it) {
                int v = it;
                if constexpr (ORDER == 1) v = (int)(xs_next(s0, s1) % (uint64_t)n);
                const int e = (v << lrpb) | slot;
                const uint8_t cur = ang[e];
                uint8_t prop;
                double pe = 0.0;
                bool flip = false;
                if constexpr (PROP == ROTOR_METROPOLIS_NE) {  // :172-190
                    prop = (uint8_t)(128 + cur);
                    const double d = dE[e];
                    pe = -d;
                    if (d <= 0.0) {
                        flip = true;
                    } else {
                        const uint64_t x = xs_next(s0, s1);
                        flip = exp(-d * beta) * kTwo64 > (double)x;
                    }
                } else if constexpr (PROP == ROTOR_GIBBS_NE) {  // :191-202
                    prop = (uint8_t)(128 + cur);
                    const double d = dE[e];
                    pe = -d;
                    const uint64_t x = xs_next(s0, s1);
                    flip = kTwo64 > (double)x * (1 + exp(d * beta));
                } else if constexpr (PROP == ROTOR_UNIFORM) {  // :203-206
                    prop = (uint8_t)(xs_next(s0, s1) % 256);
                } else {  // ROTOR_TF, :207-242
                    const uint64_t x = xs_next(s0, s1);
                    if (rat >= 1) {
                        prop = (uint8_t)(x % 256);
                    } else {
                        const int32_t osi = trunc_i32(rat * 128 - 0.5);
                        const int32_t prod = (int32_t)(uint32_t)((int64_t)osi * (int64_t)osi + (int64_t)osi);
                        const int32_t span = (int32_t)(2u * (uint32_t)osi + 1u);
                        const double distC = prod / (double)span;
                        const double p = (rat / 2 * 128 - distC) / ((int32_t)((uint32_t)osi + 1u) - distC);
                        const uint64_t y = xs_next(s0, s1);
                        uint64_t pac;
                        if (p * kTwo64 > (double)x)
                            pac = (uint64_t)(int64_t)(int32_t)((uint32_t)osi + 1u) * (2 * (y % 2) - 1);
                        else
                            pac = y % (uint64_t)(int64_t)span - (uint64_t)(int64_t)osi;
                        prop = (uint8_t)(cur + (uint8_t)pac);
                    }
                }
                if constexpr (!NE) {  // :246-266
                    double tfv;
                    if constexpr (PIPE) {
                        double f = nx.h;
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const double g = f + (nx.c[i] == vprev ? cprev : nx.t[i]) * nx.w[i];
                            f = nx.kb + i < nx.ke ? g : f;
                        }
                        pe = nx.ke - nx.kb > 8 ? field_from(f, nx.kb + 8, nx.ke) : f;
                        tfv = nx.tf;
                        fetch_ahead(v + 1 < n ? v + 1 : 0);
                    } else {
                        pe = field(v);
                        tfv = a.tf[v];
                    }
                    const double dlm = -beta * pe * (ctab[prop] - ctab[cur]) +
                                       beta_tf * tfv * (stab[prop] - stab[cur]);
                    if (dlm >= 0.0) {
                        flip = true;
                    } else {
                        const uint64_t x = xs_next(s0, s1);
                        flip = exp(dlm) * kTwo64 > (double)x;
                    }
                }
                if (flip) {  // :268-285
                    if constexpr (NE) {
                        const double mult = 2 * (ctab[cur] - ctab[prop]);
                        const uint32_t r0i = RI[v];
                        const int rb = (int)(r0i & 0x7fffffffu), re = (int)(RI[v + 1] & 0x7fffffffu);
                        if (r0i & 0x80000000u) {
                            if (j == 0)
                                for (int k = rb; k < re; ++k) {
                                    const int eu = (col_at(k) << lrpb) | slot;
                                    dE[eu] += mult * J_at(k) * ctab[ang[eu]];
                                }
                        } else {
                            for (int k = rb + j; k < re; k += L) {
                                const int eu = (col_at(k) << lrpb) | slot;
                                dE[eu] += mult * J_at(k) * ctab[ang[eu]];
                            }
                        }
                        if (j == 0) {
                            dE[e] = pe;
                            ang[e] = prop;
                        }
                    } else if (j == 0) {
                        ang[e] = prop;
                    }
                    ++accepted;
                }
                if constexpr (PIPE) {
                    vprev = v;
                    cprev = ctab[flip ? prop : cur];
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        if (snap && bi % a.interval == 0) {  // :290-294
            uint8_t* out = snap + (size_t)(bi / a.interval) * n;
            for (int v = j; v < n; v += L) out[v] = ang[(v << lrpb) | slot];
        }
    }
    __builtin_amdgcn_wave_barrier();

    // angles out and the cos-weighted energy, get_state_energyC :310-329
    for (int v = j; v < n; v += L) a.states[(size_t)r * n + v] = ang[(v << lrpb) | slot];
    if (j == 0) {
        double en = 0.0;
        for (int v = 0; v < n; ++v) en += ctab[ang[(v << lrpb) | slot]] * a.h[v];
        for (int64_t c = 0; c < a.m; ++c)
            en += ctab[ang[(a.cs[c] << lrpb) | slot]] * a.cw[c] * ctab[ang[(a.ce[c] << lrpb) | slot]];
        a.energies[r] = en;
        if (a.counters) a.counters[r] = accepted;
    }
}

// The 64-bit value of lane i (i uniform).
__device__ __forceinline__ double readlane_f64(double x, int i) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, i);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), i);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sequential order with the angle proposals (uniform, transverse-field
// window), one wave per read: the read's RNG, proposal and decision are
// wave-uniform (scalar registers), and lane k holds entry k of a row (rows of
// at most 64 entries, adjacency staged in LDS).  While one attempt decides,
// the next variable's row entries, their angles and cosines are fetched; the
// one entry the current attempt can change (a neighbour equal to it) is
// patched with the cosine of its new angle.  The local field is summed in
// adjacency order from the lanes' products, as the reference does
// (cpu_rotormc.cpp:58-79, 246-266), so every result equals rotor_kernel's.
// LDS layout as rotor_kernel's with one slot per wave.
// SUMW > 0: every row has at most SUMW entries, and the field sums lanes
// 0 .. SUMW-1 unrolled; lanes past a row's end hold -0.0, and x + (-0.0) is
// x exactly for every x, so the sum equals the reference's over the row.
template <int PROP, int SUMW>
__global__ __launch_bounds__(256) void rotor_wave_kernel(RotorArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lrpb = a.rpb_log2;
    const int n = a.n;
    const int rlocal = blockIdx.x * a.rpb + slot;
    const bool act = slot < a.rpb && rlocal < a.R;
    const int r = a.r0 + rlocal;
    double* ctab = reinterpret_cast<double*>(smem);
    double* stab = ctab + 256;
    uint32_t* RI = reinterpret_cast<uint32_t*>(smem + a.off_rowinfo);
    uint16_t* ccol = reinterpret_cast<uint16_t*>(smem + a.off_col);
    double* cJ = reinterpret_cast<double*>(smem + a.off_J);
    uint8_t* ang = smem + a.off_ang;

    for (int i = threadIdx.x; i < 512; i += blockDim.x) ctab[i] = a.tables[i];
    for (int v = threadIdx.x; v <= n; v += blockDim.x) RI[v] = (uint32_t)a.rowptr[v];
    for (int64_t k = threadIdx.x; k < a.m2; k += blockDim.x) {
        ccol[k] = (uint16_t)a.col[k];
        cJ[k] = a.J[k];
    }
    if (slot < a.rpb)
        for (int v = lane; v < n; v += 64) ang[(v << lrpb) | slot] = act ? a.states[(size_t)r * n + v] : (uint8_t)0;
    __syncthreads();
    if (!act) return;

    uint64_t s0 = a.seeds[r];
    if (s0 == 0) s0 = ~0ull;  // :407-408
    uint64_t s1 = 0;
    for (int i = 0; i < 32; ++i) (void)xs_next(s0, s1);  // burn-in, :145-147
    unsigned long long accepted = 0;
    uint8_t* snap = (a.stats && a.interval > 0) ? a.stats + (size_t)r * a.snapshots * n : nullptr;

    // The next variable's row: bounds (scalar), this lane's entry (column or
    // -1, weight, cosine of the neighbour's angle), h, t and its own angle.
    // Its loads are dependent (row bounds -> entry -> neighbour's angle ->
    // cosine); they are spread over the attempt, between the parts of the
    // decision, so their latencies overlap its work: the row bounds come one
    // attempt ahead, the entry and the variable's own angle at the top, the
    // neighbour's angle after the proposal, its cosine after the field sum.
    // Lanes past a row's end read its last entry (entry 0 of the adjacency
    // for an empty row) and leave it out.
    auto entry_of = [&](int kb, int ke) -> int {
        const int k = kb + lane;
        return k < ke ? k : (ke > 0 ? ke - 1 : 0);
    };
    int nkb = __builtin_amdgcn_readfirstlane((int)RI[0]);
    int nke = __builtin_amdgcn_readfirstlane((int)RI[1]);
    int ncol, ncur;
    double nw, nt, nh, ntf;
    {
        const int kk = entry_of(nkb, nke);
        const bool in = nkb + lane < nke;
        const int c = in ? (int)ccol[kk] : 0;
        nw = in ? cJ[kk] : 0.0;
        ncol = in ? c : -1;
        nt = ctab[ang[(c << lrpb) | slot]];
        nh = a.h[0];
        ntf = a.tf[0];
        ncur = __builtin_amdgcn_readfirstlane((int)ang[slot]);
    }
    // bounds of the row after the current one
    const int v1 = n > 1 ? 1 : 0;
    int kb1 = __builtin_amdgcn_readfirstlane((int)RI[v1]);
    int ke1 = __builtin_amdgcn_readfirstlane((int)RI[v1 + 1]);
    int vprev = -1;      // the previous attempt's variable
    double cprev = 0.0;  // cos of its angle after that attempt
    int aprev = 0;       // its angle after that attempt

    for (int bi = 0; bi < a.nf; ++bi) {
        const double beta = a.hp[bi], beta_tf = a.hd[bi];
        const double rat = beta_tf / beta;  // :157
        for (int sw = 0; sw < a.spb; ++sw) {
            for (int v = 0; v < n; ++v) {
                // this attempt's row (fetched during the previous attempt)
                const int deg = nke - nkb;
                const double pt = ncol < 0 ? -0.0 : (ncol == vprev ? cprev : nt) * nw;
                const double h = nh, tfv = ntf;
                const int cur = v == vprev ? aprev : ncur;
                // loads of the next variable's row (bounds kb1, ke1) and of
                // the bounds of the row after it
                const int vn = v + 1 < n ? v + 1 : 0;
                const int vnn = vn + 1 < n ? vn + 1 : 0;
                uint32_t rb2 = RI[vnn], re2 = RI[vnn + 1];
                const int kk = entry_of(kb1, ke1);
                const bool in = kb1 + lane < ke1;
                const int cl = (int)ccol[kk];
                double wl = cJ[kk];
                uint32_t angv = ang[(vn << lrpb) | slot];
                double hl = a.h[vn], tl = a.tf[vn];
                __builtin_amdgcn_sched_barrier(0);
                uint8_t prop;
                if constexpr (PROP == ROTOR_UNIFORM) {  // :203-206
                    prop = (uint8_t)(xs_next(s0, s1) % 256);
                } else {  // ROTOR_TF, :207-242
                    const uint64_t x = xs_next(s0, s1);
                    if (rat >= 1) {
                        prop = (uint8_t)(x % 256);
                    } else {
                        const int32_t osi = trunc_i32(rat * 128 - 0.5);
                        const int32_t prod = (int32_t)(uint32_t)((int64_t)osi * (int64_t)osi + (int64_t)osi);
                        const int32_t span = (int32_t)(2u * (uint32_t)osi + 1u);
                        const double distC = prod / (double)span;
                        const double p = (rat / 2 * 128 - distC) / ((int32_t)((uint32_t)osi + 1u) - distC);
                        const uint64_t y = xs_next(s0, s1);
                        uint64_t pac;
                        if (p * kTwo64 > (double)x)
                            pac = (uint64_t)(int64_t)(int32_t)((uint32_t)osi + 1u) * (2 * (y % 2) - 1);
                        else
                            pac = y % (uint64_t)(int64_t)span - (uint64_t)(int64_t)osi;
                        prop = (uint8_t)(cur + (uint8_t)pac);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                // the neighbour's angle (before this attempt's own update; a
                // neighbour equal to v is patched at the next attempt)
                const int c = in ? cl : 0;
                uint32_t an = ang[(c << lrpb) | slot];
                __builtin_amdgcn_sched_barrier(0);
                // the local field in adjacency order (:246-250)
                double f = h;
                if constexpr (SUMW > 0) {
#pragma unroll
                    for (int i = 0; i < SUMW; ++i) f = f + readlane_f64(pt, i);
                } else {
                    for (int i = 0; i < deg; ++i) f = f + readlane_f64(pt, i);
                }
                __builtin_amdgcn_sched_barrier(0);
                double tn = ctab[an];
                __builtin_amdgcn_sched_barrier(0);
                const double cp = ctab[prop], cc = ctab[cur];
                const double dlm = -beta * f * (cp - cc) + beta_tf * tfv * (stab[prop] - stab[cur]);
                bool flip;
                if (__builtin_amdgcn_readfirstlane(dlm >= 0.0 ? 1 : 0)) {
                    flip = true;
                } else {
                    flip = accept_exp(dlm, xs_next(s0, s1));
                }
                if (flip) {
                    if (lane == 0) ang[(v << lrpb) | slot] = prop;
                    ++accepted;
                }
                vprev = v;
                cprev = flip ? cp : cc;
                aprev = flip ? prop : cur;
                // the next attempt's row
                nkb = kb1;
                nke = ke1;
                ncol = in ? cl : -1;
                nw = in ? wl : 0.0;
                nt = tn;
                nh = hl;
                ntf = tl;
                ncur = __builtin_amdgcn_readfirstlane((int)angv);
                kb1 = __builtin_amdgcn_readfirstlane((int)rb2);
                ke1 = __builtin_amdgcn_readfirstlane((int)re2);
            }
        }
        if (snap && bi % a.interval == 0) {  // :290-294
            uint8_t* out = snap + (size_t)(bi / a.interval) * n;
            for (int v = lane; v < n; v += 64) out[v] = ang[(v << lrpb) | slot];
        }
    }
    __builtin_amdgcn_wave_barrier();

    // angles out and the cos-weighted energy, get_state_energyC :310-329
    for (int v = lane; v < n; v += 64) a.states[(size_t)r * n + v] = ang[(v << lrpb) | slot];
    if (lane == 0) {
        double en = 0.0;
        for (int v = 0; v < n; ++v) en += ctab[ang[(v << lrpb) | slot]] * a.h[v];
        for (int64_t c = 0; c < a.m; ++c)
            en += ctab[ang[(a.cs[c] << lrpb) | slot]] * a.cw[c] * ctab[ang[(a.ce[c] << lrpb) | slot]];
        a.energies[r] = en;
        if (a.counters) a.counters[r] = accepted;
    }
}

// exp(t) 2^64 > (double)x per lane (t < 0): accept_exp's test with the same
// proven margins, for decisions that differ between lanes.
__device__ __forceinline__ bool accept_exp_lane(const double t, const uint64_t x) {
    const float xh = (float)(uint32_t)(x >> 40);
    const float a = __builtin_amdgcn_exp2f((float)t * 1.44269504f) * 16777216.0f;
    if (a * 0.99993896f > xh + 1.0f) return true;
    if (a * 1.00006104f < xh) return false;
    return exp(t) * kTwo64 > (double)x;
}

// The 64-bit value of lane src (byte address src << 2 in src4).
__device__ __forceinline__ uint64_t bperm64(int src4, uint32_t lo, uint32_t hi) {
    return (uint64_t)(uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)hi) << 32 |
           (uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)lo);
}

// Sequential order, uniform angle proposals, rows of at most RW entries, one
// wave per read, in runs of at most kRotorRunMax consecutive variables no two
// of which are adjacent.  No member's field depends on another member's
// angle, so the members of a run interact only through the read's stream
// (cpu_rotormc.cpp:203-206, 252-266): member j proposes with the output after
// j + o draws, o the number of earlier members whose proposal gave a negative
// log measure and so drew an acceptance output.  Lane T_j + o (T_j =
// j (j + 1) / 2, o <= j) evaluates member j as if o such draws came first: the
// field summed in adjacency order (cpu_rotormc.cpp:58-79), the proposal, the
// log measure and, when it is negative, the acceptance against the next
// output.  Two ballots (accepted, drew) resolve the run on the scalar unit,
// member by member: o_{j+1} = o_j + drew(T_j + o_j); the chosen lanes store
// their accepted angles.  The outputs come from 64 per-lane sub-streams
// (lane l holds the read's output c + ((l - c) mod 64) of the window
// [c, c + 64)), moved on by a 64-step jump whenever fewer than
// 2 kRotorRunMax are left, as in the SA run kernel.  Every result equals
// rotor_wave_kernel's.
// EXACT: every row has exactly RW entries; otherwise rows have at most RW and
// the entries past a row's end (the staged adjacency is padded with RW zero
// entries at its end, so they stay in bounds) contribute -0.0, and x + (-0.0)
// is x exactly.  Four reads per block (slot = wave).
// PROP ROTOR_TF: the transverse-field window (:207-242) proposes with one
// output when Hd / Hp >= 1 and with two otherwise, the same for every member
// of a schedule step, so member j proposes with the outputs after
// dpp j + o draws (dpp draws per proposal) and accepts against the next.
template <int RW, bool EXACT, int PROP>
__global__ __launch_bounds__(256) void rotor_run_kernel(RotorArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int slot = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lrpb = a.rpb_log2;
    const int n = a.n;
    const int rlocal = blockIdx.x * a.rpb + slot;
    const bool act = slot < a.rpb && rlocal < a.R;
    const int r = a.r0 + rlocal;
    double* ctab = reinterpret_cast<double*>(smem);
    double* stab = ctab + 256;
    uint32_t* RI = reinterpret_cast<uint32_t*>(smem + a.off_rowinfo);
    uint16_t* ccol = reinterpret_cast<uint16_t*>(smem + a.off_col);
    double* cJ = reinterpret_cast<double*>(smem + a.off_J);
    uint8_t* ang = smem + a.off_ang;
    uint4* Jt = reinterpret_cast<uint4*>(smem + a.off_jump);

    for (int i = threadIdx.x; i < 512; i += blockDim.x) ctab[i] = a.tables[i];
    for (int v = threadIdx.x; v <= n; v += blockDim.x) RI[v] = (uint32_t)a.rowptr[v];
    for (int64_t k = threadIdx.x; k < a.m2 + RW; k += blockDim.x) {
        ccol[k] = k < a.m2 ? (uint16_t)a.col[k] : (uint16_t)0;
        cJ[k] = k < a.m2 ? a.J[k] : 0.0;
    }
    for (int t = threadIdx.x; t < kJumpEntries; t += blockDim.x) Jt[t] = reinterpret_cast<const uint4*>(a.jump)[t];
    if (slot < a.rpb)
        for (int v = lane; v < n; v += 64) ang[(v << lrpb) | slot] = act ? a.states[(size_t)r * n + v] : (uint8_t)0;
    __syncthreads();
    if (!act) return;

    uint8_t* angs = ang + slot;  // this read's angle of v at angs[4 v]
    // this lane's pair: member j of a run, o acceptance draws before it
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= lane) ++j;
    const int o = lane - j * (j + 1) / 2;

    // ---- the read's stream after its burn-in (:145-147); lane l first holds
    // output l, its sub-stream state before that output in (q0, q1, q2, q3)
    uint64_t s0 = a.seeds[r];
    s0 = ((uint64_t)__builtin_amdgcn_readfirstlane((int)(s0 >> 32)) << 32) |
         (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)s0);
    if (s0 == 0) s0 = ~0ull;  // :407-408
    uint64_t s1 = 0;
    for (int i = 0; i < 32; ++i) (void)xs_next(s0, s1);
    uint32_t q0 = (uint32_t)s0, q1 = (uint32_t)(s0 >> 32), q2 = (uint32_t)s1, q3 = (uint32_t)(s1 >> 32);
    for (int i = 0; i < 63; ++i) {
        const XsStep st = xs_peek(q0, q1, q2, q3);
        const bool go = i < lane;
        q0 = go ? q2 : q0;
        q1 = go ? q3 : q1;
        q2 = go ? st.nlo : q2;
        q3 = go ? st.nhi : q3;
    }
    uint32_t wlo, whi;
    {
        const XsStep st = xs_peek(q0, q1, q2, q3);
        wlo = st.olo;
        whi = st.ohi;
    }
    uint32_t held = (uint32_t)lane, cpos = 0, wend = 64;  // held: the output this lane holds

    unsigned long long accepted = 0;
    uint8_t* snap = (a.stats && a.interval > 0) ? a.stats + (size_t)r * a.snapshots * n : nullptr;
    const int nruns = a.nruns;
    // What a run needs that no run changes, loaded ahead (the runs repeat
    // every sweep): the member table entry of this lane's slot (row start;
    // member | degree << 16 | run length << 24) two runs ahead, then the
    // member's row (columns, weights), h and transverse field one run ahead.
    // The loads issued during a run depend on nothing loaded in it.
    const uint2* mtab = reinterpret_cast<const uint2*>(a.run_rows) + (j < 16 ? j : 15);
    auto next_run = [&](int ri) { return ri + 1 < nruns ? ri + 1 : 0; };
    uint2 t0 = mtab[0], t1 = mtab[16 * next_run(0)];
    uint16_t c0[RW];
    double w0[RW];
    auto load_row = [&](const uint2 t, uint16_t (&c)[RW], double (&w)[RW]) {
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            c[i] = ccol[t.x + i];
            w[i] = cJ[t.x + i];
        }
    };
    load_row(t0, c0, w0);
    double h0 = a.h[t0.y & 0xffffu], f0 = a.tf[t0.y & 0xffffu];

    for (int bi = 0; bi < a.nf; ++bi) {
        const double beta = a.hp[bi], beta_tf = a.hd[bi];
        // the transverse-field window of this step (:157, :207-229), as the
        // wave kernel computes it per attempt
        const double rat = beta_tf / beta;
        const bool wide = PROP != ROTOR_TF || rat >= 1;  // one output per proposal
        const int dpp = wide ? 1 : 2;
        int32_t osi = 0, span = 1;
        double pw = 0.0;
        if (!wide) {
            osi = trunc_i32(rat * 128 - 0.5);
            const int32_t prod = (int32_t)(uint32_t)((int64_t)osi * (int64_t)osi + (int64_t)osi);
            span = (int32_t)(2u * (uint32_t)osi + 1u);
            const double distC = prod / (double)span;
            pw = (rat / 2 * 128 - distC) / ((int32_t)((uint32_t)osi + 1u) - distC);
        }
        for (int sw = 0; sw < a.spb; ++sw) {
            for (int ri = 0; ri < nruns; ++ri) {
                // this run's entry, row, h and t; the loads for the next two
                const int v = (int)(t0.y & 0xffffu);
                const int deg = EXACT ? RW : (int)((t0.y >> 16) & 0xffu);
                const int K = __builtin_amdgcn_readfirstlane((int)(t0.y >> 24));
                const double hv = h0, tfv = f0;
                uint16_t rc[RW];
                double rJ[RW];
#pragma unroll
                for (int i = 0; i < RW; ++i) {
                    rc[i] = c0[i];
                    rJ[i] = w0[i];
                }
                const uint2 t2 = mtab[16 * next_run(next_run(ri))];
                load_row(t1, c0, w0);
                h0 = a.h[t1.y & 0xffffu];
                f0 = a.tf[t1.y & 0xffffu];
                t0 = t1;
                t1 = t2;
                const bool live = j < K;
                // the proposal's outputs and the one after them
                const uint32_t p = cpos + (uint32_t)(dpp * j + o);
                const uint64_t x1 = bperm64((int)(p & 63u) << 2, wlo, whi);
                const uint64_t x2 = bperm64((int)((p + 1u) & 63u) << 2, wlo, whi);
                uint64_t xa = x2;  // the acceptance output
                if (PROP == ROTOR_TF && !wide) xa = bperm64((int)((p + 2u) & 63u) << 2, wlo, whi);
                // the field in adjacency order (:58-79, 246-250)
                double f = hv;
                double pt[RW];
#pragma unroll
                for (int i = 0; i < RW; ++i) {
                    const double c = ctab[angs[4 * (int)rc[i]]] * rJ[i];
                    pt[i] = EXACT || i < deg ? c : -0.0;
                }
#pragma unroll
                for (int i = 0; i < RW; ++i) f = f + pt[i];
                const int cur = angs[4 * v];
                int prop = (int)(x1 & 255u);  // :204-206, :212-213
                if (PROP == ROTOR_TF && !wide) {  // :233-238
                    uint64_t pac;
                    if (pw * kTwo64 > (double)x1)
                        pac = (uint64_t)(int64_t)(int32_t)((uint32_t)osi + 1u) * (2 * (x2 % 2) - 1);
                    else
                        pac = x2 % (uint64_t)(int64_t)span - (uint64_t)(int64_t)osi;
                    prop = (uint8_t)(cur + (uint8_t)pac);
                }
                const double cp = ctab[prop], cc = ctab[cur];
                const double dlm = -beta * f * (cp - cc) + beta_tf * tfv * (stab[prop] - stab[cur]);
                const bool drew = !(dlm >= 0.0);
                const bool acc = drew ? accept_exp_lane(dlm, xa) : true;  // :256-266
                const uint64_t A = __builtin_amdgcn_ballot_w64(live && acc);
                const uint64_t C = __builtin_amdgcn_ballot_w64(live && drew);
                // resolve: member jj's lane is T_jj + (acceptance draws before
                // it); straight-line over kRotorRunMax members (members past
                // the run have no bits in A or C, so they add no draws and
                // choose no accepted lane)
                uint32_t od = 0;
                uint64_t chosen = 0;
#pragma unroll
                for (int jj = 0; jj < kRotorRunMax; ++jj) {
                    const int L = jj * (jj + 1) / 2 + (int)od;
                    chosen |= 1ull << L;
                    od += (uint32_t)(C >> L) & 1u;
                }
                const uint64_t took = chosen & A;
                if ((took >> lane) & 1ull) angs[4 * v] = (uint8_t)prop;
                accepted += (unsigned long long)__popcll(took);
                // ---- move the window on by the run's dpp K + od outputs ------
                cpos += (uint32_t)(dpp * K) + od;
                if (wend - cpos < (PROP == ROTOR_TF ? 3u : 2u) * kRotorRunMax) {
                    const bool stale = (int32_t)(held - cpos) < 0;
                    uint32_t x0 = q0, xq1 = q1, y0 = q2, y1 = q3;
                    xs_jump_lds(x0, xq1, y0, y1, Jt);
                    const XsStep st = xs_peek(x0, xq1, y0, y1);
                    q0 = stale ? x0 : q0;
                    q1 = stale ? xq1 : q1;
                    q2 = stale ? y0 : q2;
                    q3 = stale ? y1 : q3;
                    wlo = stale ? st.olo : wlo;
                    whi = stale ? st.ohi : whi;
                    held = stale ? held + 64u : held;
                    wend = cpos + 64u;
                }
            }
        }
        if (snap && bi % a.interval == 0) {  // :290-294
            uint8_t* out = snap + (size_t)(bi / a.interval) * n;
            for (int v = lane; v < n; v += 64) out[v] = ang[(v << lrpb) | slot];
        }
    }
    __builtin_amdgcn_wave_barrier();

    // angles out and the cos-weighted energy, get_state_energyC :310-329
    for (int v = lane; v < n; v += 64) a.states[(size_t)r * n + v] = ang[(v << lrpb) | slot];
    if (lane == 0) {
        double en = 0.0;
        for (int v = 0; v < n; ++v) en += ctab[ang[(v << lrpb) | slot]] * a.h[v];
        for (int64_t c = 0; c < a.m; ++c)
            en += ctab[ang[(a.cs[c] << lrpb) | slot]] * a.cw[c] * ctab[ang[(a.ce[c] << lrpb) | slot]];
        a.energies[r] = en;
        if (a.counters) a.counters[r] = accepted;
    }
}

template <int P>
auto run_variant(int row_max, bool ex) {
    return row_max <= 4   ? (ex && row_max == 4 ? rotor_run_kernel<4, true, P> : rotor_run_kernel<4, false, P>)
           : row_max <= 6 ? (ex && row_max == 6 ? rotor_run_kernel<6, true, P> : rotor_run_kernel<6, false, P>)
           : row_max <= 8 ? (ex && row_max == 8 ? rotor_run_kernel<8, true, P> : rotor_run_kernel<8, false, P>)
                          : (ex && row_max == 16 ? rotor_run_kernel<16, true, P> : rotor_run_kernel<16, false, P>);
}

template <int PROP>
hipError_t launch_wave(const RotorArgs& a, int nblocks, hipStream_t stream) {
    auto fn = a.row_max <= 8 ? rotor_wave_kernel<PROP, 8>
                             : (a.row_max <= 16 ? rotor_wave_kernel<PROP, 16> : rotor_wave_kernel<PROP, 0>);
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(a.threads), a.lds_bytes, stream, a);
    return hipGetLastError();
}

template <int ORDER, int PROP, bool CSR>
hipError_t launch_variant(const RotorArgs& a, int nblocks, hipStream_t stream) {
    auto fn = rotor_kernel<ORDER, PROP, CSR>;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(a.threads), a.lds_bytes, stream, a);
    return hipGetLastError();
}

template <int ORDER, int PROP>
hipError_t launch_csr(const RotorArgs& a, int nblocks, hipStream_t stream) {
    return a.csr_in_lds ? launch_variant<ORDER, PROP, true>(a, nblocks, stream)
                        : launch_variant<ORDER, PROP, false>(a, nblocks, stream);
}

template <int ORDER>
hipError_t launch_prop(const RotorArgs& a, int proposal, int nblocks, hipStream_t stream) {
    switch (proposal) {
        case ROTOR_GIBBS_NE: return launch_csr<ORDER, ROTOR_GIBBS_NE>(a, nblocks, stream);
        case ROTOR_METROPOLIS_NE: return launch_csr<ORDER, ROTOR_METROPOLIS_NE>(a, nblocks, stream);
        case ROTOR_UNIFORM: return launch_csr<ORDER, ROTOR_UNIFORM>(a, nblocks, stream);
        default: return launch_csr<ORDER, ROTOR_TF>(a, nblocks, stream);
    }
}

}  // namespace

hipError_t launch_rotor(const RotorArgs& a, int randomize, int proposal, int nblocks, hipStream_t stream) {
    if (nblocks <= 0 || a.R <= 0) return hipSuccess;
    if (a.run_kernel && !randomize && (proposal == ROTOR_UNIFORM || proposal == ROTOR_TF)) {
        const bool ex = a.row_min == a.row_max;
        auto fn = proposal == ROTOR_UNIFORM ? run_variant<ROTOR_UNIFORM>(a.row_max, ex)
                                            : run_variant<ROTOR_TF>(a.row_max, ex);
        if (a.lds_bytes > 65536) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(fn, dim3(nblocks), dim3(a.threads), a.lds_bytes, stream, a);
        return hipGetLastError();
    }
    if (a.wave_per_read && !randomize && (proposal == ROTOR_UNIFORM || proposal == ROTOR_TF))
        return proposal == ROTOR_UNIFORM ? launch_wave<ROTOR_UNIFORM>(a, nblocks, stream)
                                         : launch_wave<ROTOR_TF>(a, nblocks, stream);
    return randomize ? launch_prop<1>(a, proposal, nblocks, stream) : launch_prop<0>(a, proposal, nblocks, stream);
}

}  // namespace misa
