// xorshift128+ on the device (the reference's generator,
// dwave/samplers/sa/src/cpu_sa.cpp:27-34), in the forms the integer sweep
// kernels use: a plain step, and a step on 32-bit halves that returns the
// new state word and the output without committing them.
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

}  // namespace misa
