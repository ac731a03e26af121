// Device helpers shared by the front kernels: the assembly of a dense front held in
// LDS (A through the owner-grouped assembly runs, then the children's contribution
// blocks in child_list order), in the order the CPU reference uses
// (src/reference/cholesky.jl:73-92, src/reference/ldlt.jl:176-190): every front entry
// receives its A entries first, summed in nzval order, then each child's value in turn.
#pragma once
#include "sds_common.h"

namespace sds {

// F (f x f, column-major, any address space) <- 0 + A + children; `full`: the children's
// whole m x m blocks (LU), else their lower triangles. Every thread of the block calls it;
// it ends with a barrier.
template <typename T>
__device__ void assemble_dense(const DevSym& S, int s, const double* __restrict__ nz, const T* __restrict__ cbpool,
                               T* F, int f, int full) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int q = tid; q < f * f; q += nt) F[q] = T(0);
    __syncthreads();
    // one thread sums each run of entries with the same destination, in nzval order
    const int64_t a0 = S.asm_ptr[s], a1 = S.asm_ptr[s + 1];
    for (int64_t k = a0 + tid; k < a1; k += nt) {
        const int64_t d = S.asm_dst[k];
        if (k > a0 && S.asm_dst[k - 1] == d) continue;
        double v = 0.0;
        for (int64_t q = k; q < a1 && S.asm_dst[q] == d; ++q) v += nz[S.asm_src[q]];
        if (d >= 0) {
            F[d] += T(v);
        } else {
            // a mirrored entry: the U^T structure of an unsymmetric front is stored transposed
            const int64_t loc = -(d + 1);
            if (S.unsym) {
                const int64_t j = loc / f, i = loc - j * f;
                F[i * f + j] += T(v);
            } else {
                F[loc] += T(v);
            }
        }
    }
    __syncthreads();
    for (int k = S.child_ptr[s]; k < S.child_ptr[s + 1]; ++k) {
        const int c = S.child_list[k];
        const int64_t co = S.cb_off[c];
        if (co < 0) continue;
        const int mc = S.fnrows[c] - S.fncols[c];
        const T* C = cbpool + co;
        const int32_t* ri = S.relind + S.relind_ptr[c];
        for (int q = tid; q < mc * mc; q += nt) {
            const int jj = q / mc, ii = q - jj * mc;
            if (!full && ii < jj) continue;
            F[ri[jj] * f + ri[ii]] += C[q];
        }
        __syncthreads();
    }
}

}  // namespace sds
