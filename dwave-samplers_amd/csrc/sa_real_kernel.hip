// Real-valued (float64) simulated-annealing sweep kernel for gfx950.
//
// Same algorithm, float64 operation order and RNG stream as the reference
// core (dwave/samplers/sa/src/cpu_sa.cpp:93-197, xorshift128+ :27-34).  Per
// read the kernel keeps the local field f_v (float64) and the spin s_v (int8);
// the reference's dE_v equals -2 s_v f_v exactly (scaling by -2 s_v is exact),
// and an accepted flip of v adds -2 J s_v to each neighbour's field, which
// rounds exactly like the reference's dE_u += 4 s_v J s_u.
//
// Block = waves_per_block waves; each wave holds 64/L reads, L lanes per read.
// LDS regions (byte offsets in SaArgs):
//   [0, off_S)          fields, variable-major: f(v, slot) at (v << log2 rpb) | slot
//   [off_S, off_rowinfo) spins, same indexing
//   off_rowinfo         rowptr[v] | (serial << 31) | (one entry to v+1 << 30), n+1 entries
//   off_col, off_J      CSR column indices (u16) and weights (f64), when staged
// A row is serial when it repeats a neighbour or holds v itself: its updates
// are applied by one lane in CSR order (the reference's float order); other
// rows update their distinct neighbours in parallel with return-less float64
// LDS atomics.  Acceptance evaluates the reference's double expressions with
// the device exp.
//
// Sequential order (n >= 2; with the adjacency in LDS, rows no longer than
// the lanes of a read) runs a pipelined loop: the row info of v+3,
// this lane's CSR entry and J(w, w+1) of row w = v+2 (the entry from LDS or
// device memory, J from device memory), and the field and spin of v+1 are all
// loaded at the top of step v; the decision of v then uses registers only,
// and the loads land after it.  The field of v+1 was read before v's
// updates, so v's flip is applied to the register copy: + (-2 s_v) J(v, v+1)
// when exactly one row-v entry holds v+1 (the same product and sum as the
// update itself), a fresh read after a serial row's ordered updates.
#include "sa_real_block.h"

namespace misa {

template <int ORDER, int PROP, bool CSR_LDS>
__global__ __launch_bounds__(256) void sa_real_kernel(SaArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    real_block<ORDER, PROP, CSR_LDS>(a, blockIdx.x, smem);
}

template <int ORDER, int PROP, bool CSR>
static hipError_t launch_real_variant(const SaArgs& a, int nblocks, hipStream_t stream) {
    auto fn = sa_real_kernel<ORDER, PROP, CSR>;
    if (a.lds_bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)a.lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(64 * a.waves_per_block), a.lds_bytes, stream, a);
    return hipGetLastError();
}

template <int ORDER, int PROP>
static hipError_t launch_real_csr(const SaArgs& a, bool csr, int nb, hipStream_t s) {
    return csr ? launch_real_variant<ORDER, PROP, true>(a, nb, s) : launch_real_variant<ORDER, PROP, false>(a, nb, s);
}

hipError_t launch_sa_real_lds(const SaArgs& a, int order, int proposal, bool csr_in_lds, int nblocks,
                              hipStream_t stream) {
    if (nblocks <= 0) return hipSuccess;
    if (order == 0)
        return proposal == 1 ? launch_real_csr<0, 1>(a, csr_in_lds, nblocks, stream)
                             : launch_real_csr<0, 0>(a, csr_in_lds, nblocks, stream);
    return proposal == 1 ? launch_real_csr<1, 1>(a, csr_in_lds, nblocks, stream)
                         : launch_real_csr<1, 0>(a, csr_in_lds, nblocks, stream);
}

}  // namespace misa
