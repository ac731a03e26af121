// Dense front kernels' device helpers: the Cholesky of a front held in LDS and the inverse of
// its diagonal block, shared by the per-level small-front kernel and the subtree kernel.
#pragma once
#include "sds_common.h"

namespace sds {

// Cholesky pivot acceptance on the FP64 value of the pivot (src/numeric/front.jl:46-50):
// a pivot in [0, eps] is raised to eps (cuDSS pivot_epsilon), negative or NaN fails.
__device__ inline bool chol_accept(double a, double eps) { return a > eps || (eps > 0.0 && a >= 0.0); }

// panel width of the blocked dense Cholesky
constexpr int DPB = 16;

// LDS written by some lanes of a wave is read by others after this (in-order LDS per wave; the
// fence keeps the compiler from moving accesses across it and waits for the outstanding ones)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of v in lane `lane` (wave-uniform) with v_readlane (no LDS round trip)
__device__ inline double lane_value(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ inline float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Right-looking Cholesky of the first w columns of the dense f x f front F (LDS, column-major,
// ld f; f <= 64 RPL) by a workgroup of 256 threads, in panels of DPB columns:
//  - wave 0 factors the panel in registers (rows over the lanes, RPL rows per lane; pivots and
//    the panel's rows broadcast with lane reads): column j gets d = sqrt(max(a_jj, eps)) and the
//    rows below are multiplied by 1/d (as the CPU reference's potf2 scaling), then each later
//    panel column c gets -= L(i, j) L(c, j);
//  - all four waves give the columns right of the panel (the rest of the fully summed columns and
//    the contribution block, lower triangle) the panel's update, a wave per column, rows over the
//    lanes, the panel's rows of the lane held in registers.
// Two barriers per panel. Every entry receives its column updates one by one in column order, as
// a separate product and difference, exactly as in the column-by-column form. Returns 0, or the
// local column (1-based) of the first failed pivot (the columns from it on are left unfactored).
// `sh` is one int of LDS. Every thread of the block calls it; it ends with a barrier.
template <typename T, int RPL>
__device__ int front_chol_lds(T* F, int f, int w, double eps, int* sh) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) *sh = 0;
    __syncthreads();
    for (int k0 = 0; k0 < w; k0 += DPB) {
        const int kb = min(DPB, w - k0), k1 = k0 + kb;
        if (wv == 0) {
            T p[RPL][DPB];
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                const int i = k0 + lane + 64 * q;
#pragma unroll
                for (int c = 0; c < DPB; ++c) p[q][c] = (i < f && c < kb) ? F[(k0 + c) * f + i] : T(0);
            }
            int st = 0;
#pragma unroll
            for (int c = 0; c < DPB; ++c) {
                if (c < kb && !st) {
                    const int j = k0 + c;
                    const double a = (double)lane_value(p[0][c], c);
                    if (!chol_accept(a, eps)) {
                        st = j + 1;
                    } else {
                        const T d = T(sqrt(fmax(a, eps)));
                        const T rd = T(1) / d;
#pragma unroll
                        for (int q = 0; q < RPL; ++q) {
                            const int i = k0 + lane + 64 * q;
                            if (i == j) p[q][c] = d;
                            else if (i > j) p[q][c] = p[q][c] * rd;
                        }
#pragma unroll
                        for (int c2 = c + 1; c2 < DPB; ++c2) {
                            if (c2 < kb) {
                                const T l = lane_value(p[0][c], c2);
#pragma unroll
                                for (int q = 0; q < RPL; ++q) {
                                    const int i = k0 + lane + 64 * q;
                                    if (i >= k0 + c2) {
                                        const T prod = p[q][c] * l;
                                        p[q][c2] = p[q][c2] - prod;
                                    }
                                }
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                const int i = k0 + lane + 64 * q;
#pragma unroll
                for (int c = 0; c < DPB; ++c)
                    if (i < f && c < kb && i >= k0 + c) F[(k0 + c) * f + i] = p[q][c];
            }
            if (lane == 0 && st) *sh = st;
        }
        __syncthreads();
        const int st = *sh;
        if (st) return st;
        // the panel's update of the columns right of it, lower triangle
        T pr[RPL][DPB];
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            const int i = k1 + lane + 64 * q;
#pragma unroll
            for (int c = 0; c < DPB; ++c) pr[q][c] = (i < f && c < kb) ? F[(k0 + c) * f + i] : T(0);
        }
        for (int j = k1 + wv; j < f; j += 4) {
            T lj[DPB];
#pragma unroll
            for (int c = 0; c < DPB; ++c) lj[c] = c < kb ? F[(k0 + c) * f + j] : T(0);
#pragma unroll
            for (int q = 0; q < RPL; ++q) {
                const int i = k1 + lane + 64 * q;
                if (i >= j && i < f) {
                    T v = F[j * f + i];
#pragma unroll
                    for (int c = 0; c < DPB; ++c)
                        if (c < kb) {
                            const T prod = pr[q][c] * lj[c];
                            v = v - prod;
                        }
                    F[j * f + i] = v;
                }
            }
        }
        __syncthreads();
    }
    return 0;
}

// Inverse of the lower-triangular w x w diagonal block L11 of the factored front F (LDS, ld f),
// w <= BS <= NB, into `out` (column-major, ld NB: column c at out + c NB, rows past w zero):
// thread c < w computes column c by forward substitution, x = L11^-1 e_c (x_m *= 1 / L_mm, then
// x_r -= L_rm x_m), the arithmetic of the per-block inverse kernel.
template <typename T, int BS>
__device__ void diag_inverse_bs(const T* F, int f, int w, T* __restrict__ out) {
    // column c of the block is factor column super_ptr[s] + c: only the front's own w columns
    const int c = threadIdx.x;
    if (c >= w) return;
    T x[BS];
#pragma unroll
    for (int r = 0; r < BS; ++r) x[r] = r == c ? T(1) : T(0);
#pragma unroll
    for (int m = 0; m < BS; ++m) {
        if (m < w) {
            const T rinv = T(1) / F[m * f + m];
            x[m] = x[m] * rinv;
            const T xm = x[m];
#pragma unroll
            for (int r = m + 1; r < BS; ++r)
                if (r < w) x[r] = fma(-F[m * f + r], xm, x[r]);
        }
    }
    T* o = out + (int64_t)c * NB;
#pragma unroll
    for (int r = 0; r < BS; ++r) o[r] = r < w ? x[r] : T(0);
#pragma unroll
    for (int r = BS; r < NB; ++r) o[r] = T(0);
}

template <typename T>
__device__ void diag_inverse_lds(const T* F, int f, int w, T* __restrict__ out) {
    diag_inverse_bs<T, FUSED_INV_MAX>(F, f, w, out);
}

}  // namespace sds
