// xorshift128+ on the device (the reference's generator,
// dwave/samplers/sa/src/cpu_sa.cpp:27-34), in the forms the integer sweep
// kernels use: a plain step, a step on 32-bit halves that returns the new
// state word and the output without committing them, and a jump of many
// steps through a table in LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

__device__ __forceinline__ uint64_t xs_next(uint64_t& s0, uint64_t& s1) {
    uint64_t a = s0;
    const uint64_t b = s1;
    s0 = b;
    a ^= a << 23;
    s1 = a ^ b ^ (a >> 17) ^ (b >> 26);
    return s1 + b;
}

// x ^ y ^ z in one v_bitop3_b32 (truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t x, uint32_t y, uint32_t z) {
    return __builtin_amdgcn_bitop3_b32(x, y, z, 0x96);
}

// One step from the state (s0, s1) = (a, b): the new s1 word n and the
// output n + b (the new state is (b, n)).
struct XsStep {
    uint32_t nlo, nhi;   // new s1
    uint32_t olo, ohi;   // output n + s1
};

__device__ __forceinline__ XsStep xs_peek(uint32_t alo, uint32_t ahi, uint32_t blo, uint32_t bhi) {
    // a ^= a << 23
    const uint32_t tlo = alo << 23;
    const uint32_t thi = __builtin_amdgcn_alignbit(ahi, alo, 9);
    alo ^= tlo;
    ahi ^= thi;
    // n = a ^ b ^ (a >> 17) ^ (b >> 26)
    XsStep r;
    r.nlo = xor3(alo, blo, __builtin_amdgcn_alignbit(ahi, alo, 17)) ^ __builtin_amdgcn_alignbit(bhi, blo, 26);
    r.nhi = xor3(ahi, bhi, ahi >> 17) ^ (bhi >> 26);
    const uint64_t o = ((uint64_t)r.nhi << 32 | r.nlo) + ((uint64_t)bhi << 32 | blo);
    r.olo = (uint32_t)o;
    r.ohi = (uint32_t)(o >> 32);
    return r;
}

// (x0, x1, y0, y1) = (s0, s1) as 32-bit halves, advanced by the table's
// number of xorshift128+ steps: the state update is linear over GF(2), so the result is
// the XOR over the 32 nibbles of the 128-bit state (bit 4c + j is bit j of
// nibble c; s0 holds bits 0..63) of tab[16 c + nibble] (the advanced images
// of the nibble's bits, built on the host).
__device__ __forceinline__ void xs_jump(uint32_t& x0, uint32_t& x1, uint32_t& y0, uint32_t& y1, const uint4* tab) {
    const uint32_t w[4] = {x0, x1, y0, y1};
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    // eight table loads in flight at a time, then their XOR
#pragma unroll
    for (int c0 = 0; c0 < 32; c0 += 8) {
        uint4 e[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) e[c] = tab[(c0 + c) * 16 + ((w[(c0 + c) >> 3] >> (4 * c)) & 15u)];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < 8; c += 2) {
            r0 = xor3(r0, e[c].x, e[c + 1].x);
            r1 = xor3(r1, e[c].y, e[c + 1].y);
            r2 = xor3(r2, e[c].z, e[c + 1].z);
            r3 = xor3(r3, e[c].w, e[c + 1].w);
        }
    }
    x0 = r0;
    x1 = r1;
    y0 = r2;
    y1 = r3;
}

}  // namespace misa
