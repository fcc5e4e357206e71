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

// base + byte K of w, zero-extended: one SDWA add.
template <int K>
__device__ __forceinline__ uint32_t add_byte(uint32_t base, uint32_t w) {
    uint32_t r;
    if constexpr (K == 0)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
            : "=v"(r) : "v"(base), "v"(w));
    else if constexpr (K == 1)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1"
            : "=v"(r) : "v"(base), "v"(w));
    else if constexpr (K == 2)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2"
            : "=v"(r) : "v"(base), "v"(w));
    else
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3"
            : "=v"(r) : "v"(base), "v"(w));
    return r;
}

// xs_jump with the table in LDS, each entry's address formed by one SDWA add
// (used where the jump is a measured share of the time: the rotor run kernel).
__device__ __forceinline__ void xs_jump_lds(uint32_t& x0, uint32_t& x1, uint32_t& y0, uint32_t& y1, const uint4* tab) {
    const uint32_t w[4] = {x0, x1, y0, y1};
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    // the table's LDS byte address; entry 16 c + v of nibble c of word k
    // (16 bytes each) lies at byte 256 c plus byte c / 2 of the word's even
    // nibbles shifted up by 4 (c even) or of its odd nibbles in place (c odd)
    const uint32_t tb = (uint32_t)(size_t)(const __attribute__((address_space(3))) uint4*)tab;
    // eight table loads in flight at a time, then their XOR
#pragma unroll
    for (int c0 = 0; c0 < 32; c0 += 8) {
        const uint32_t we = (w[c0 >> 3] & 0x0f0f0f0fu) << 4, wo = w[c0 >> 3] & 0xf0f0f0f0u;
        uint32_t ad[8];
        ad[0] = add_byte<0>(tb, we);
        ad[1] = add_byte<0>(tb, wo);
        ad[2] = add_byte<1>(tb, we);
        ad[3] = add_byte<1>(tb, wo);
        ad[4] = add_byte<2>(tb, we);
        ad[5] = add_byte<2>(tb, wo);
        ad[6] = add_byte<3>(tb, we);
        ad[7] = add_byte<3>(tb, wo);
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        u32x4 e[8];
#pragma unroll
        for (int c = 0; c < 8; ++c)
            e[c] = *reinterpret_cast<const __attribute__((address_space(3))) u32x4*>(
                (size_t)(ad[c] + (uint32_t)(c0 + c) * 256u));
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
