// Cholesky sweeps with inverted diagonal blocks (gfx950).
//
// After a Cholesky factorization one launch inverts every NB x NB diagonal block of every
// front (Linv = L_bb^-1, lower triangular, kept in the DB buffer: NB entries per factor column,
// column-major with leading dimension NB). The sweeps then need no serial substitution: a
// front's fully summed part is a short sequence of NB-row GEMVs done by one workgroup, and the
// tall L21 products of the big fronts are separate, fully parallel launches. This is the
// partitioned-inverse form of the reference's solve_alg = "algo1" (src/solver.jl:834,
// src/solve/fused.jl) applied per diagonal block.
//
// Per tree level, for the narrow fronts (w <= NB, one diagonal block):
//   forward : k_fwd_front (one workgroup per front: pull the children's update vectors, solve
//             the block, and for fronts with few rows below, their L21 update too), then
//             k_fwd_l21 for the tall ones (row tiles of L21 times the solved top);
//   backward: k_bwd_l21t for the tall ones (column blocks of L21^T times the gathered
//             solution, partial sums per row chunk), then k_bwd_front (sum the partials in
//             chunk order, or form L21^T x itself, and solve the block).
// The wide fronts (w > NB) run their blocks as level-batched steps, many workgroups per front:
//   forward : k_fwd_front pulls the children's vectors, then per block k one k_fwd_wstep launch
//             (every workgroup forms y_k = Linv_kk y_k and applies it to its tile of 256 rows
//             below; the task with c = 1 keeps y_k in Z), then the solved blocks Z -> Y;
//   backward: k_bwd_l21t partials, k_bwd_front sums them into the top rows, then per block from
//             the last one k_bwd_wstep launch (every workgroup forms x_k = Linv_kk^T y_k and
//             subtracts its contribution from a tile of 64 earlier columns; c = 1 keeps x_k in
//             Z), then Z -> Y.
// Every entry is produced by one thread with a fixed summation order: no atomics, and the
// result is bitwise reproducible.
#include "sds_common.h"
#include "sds_kernels.h"

namespace sds {

// ---------------------------------------------------------------------------
// Linv of the diagonal block [k, k + nb) of front s: one wave; lane c computes column c of
// the inverse by forward substitution with the partial sums in registers (fully unrolled:
// the block is padded with the identity past nb, so no guards touch the register array;
// no scheduling barriers either: with them the compiler spills this kernel). The block is
// read from the panel (small fronts) or, for the large fronts (c = 1), from DB, where the
// block steps left it; it is then also written into its panel. Linv replaces it in DB.
// task: (s, k, 0, large)

template <typename T>
__global__ void __launch_bounds__(64) k_block_inverse(DevSym S, const Task* __restrict__ tasks,
                                                     T* __restrict__ factor, T* __restrict__ DB) {
    __shared__ T Lb[NB][NB + 1];  // Lb[m][r] = L(k + r, k + m)
    __shared__ T rinv[NB];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a;
    const int nb = min(NB, w - k);
    T* P = factor + S.panel_off[s];
    T* blk = DB + (int64_t)(S.super_ptr[s] + k) * NB;
    const int c = threadIdx.x;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int r = c;
        T v = r == m ? T(1) : T(0);
        if (m < nb && r < nb) v = r >= m ? (t.c ? blk[(int64_t)m * NB + r] : P[(int64_t)(k + m) * f + k + r]) : T(0);
        Lb[m][r] = v;
    }
    __syncthreads();
    if (t.c && c < nb)
        for (int m = 0; m <= c; ++m) P[(int64_t)(k + m) * f + k + c] = Lb[m][c];
    rinv[c] = T(1) / Lb[c][c];
    __syncthreads();
    T x[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) x[r] = r == c ? T(1) : T(0);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        x[m] = x[m] * rinv[m];
        const T xm = x[m];
#pragma unroll
        for (int r = m + 1; r < NB; ++r) x[r] = fma(-Lb[m][r], xm, x[r]);
    }
    // column c of the block (only the front's own columns: the last block may be narrow)
    T* out = blk;
    if (c < nb)
#pragma unroll
        for (int r = 0; r < NB; ++r) out[(int64_t)c * NB + r] = r < nb ? x[r] : T(0);
}

// ---------------------------------------------------------------------------
// helpers

// ys[j * nrhs + r] (j < nb) = sum_{i < nb} Linv[j][i] xs[i * nrhs + r] (trans = 0), or
// sum_i Linv[i][j] xs[i] (trans = 1); Linv column-major ld NB in global memory (L2 resident).
template <typename T>
__device__ inline void block_gemv(const T* __restrict__ Li, const T* xs, T* ys, int nb, int nrhs, int trans) {
    for (int q = threadIdx.x; q < nb * nrhs; q += blockDim.x) {
        const int j = q / nrhs, r = q - j * nrhs;
        T acc = T(0);
        if (!trans) {
            for (int i = 0; i <= j; ++i) acc += Li[(int64_t)i * NB + j] * xs[i * nrhs + r];
        } else {
            for (int i = j; i < nb; ++i) acc += Li[(int64_t)j * NB + i] * xs[i * nrhs + r];
        }
        ys[q] = acc;
    }
}

constexpr int L21_ROWS = 256;  // rows of L21 per workgroup of k_fwd_l21

// ---------------------------------------------------------------------------
// forward

// one workgroup per front of the level. fronts: the level's fronts
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_front(DevSym S, const int32_t* __restrict__ fronts,
                                                  const T* __restrict__ factor, const T* __restrict__ DB,
                                                  T* __restrict__ Y, T* __restrict__ U, int nrhs) {
    __shared__ T xs[NB * MAX_NRHS];
    __shared__ T ys[NB * MAX_NRHS];
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const T* P = factor + S.panel_off[s];
    T* us = S.uv_off[s] >= 0 ? U + S.uv_off[s] * nrhs : nullptr;
    if (us)
        for (int q = tid; q < m * nrhs; q += nt) us[q] = T(0);
    __syncthreads();
    // the children's update vectors, in child_list order
    for (int kc = S.child_ptr[s]; kc < S.child_ptr[s + 1]; ++kc) {
        const int c = S.child_list[kc];
        if (S.uv_off[c] < 0) continue;
        const int mc = S.fnrows[c] - S.fncols[c];
        const T* uc = U + S.uv_off[c] * nrhs;
        const int32_t* ri = S.relind + S.relind_ptr[c];
        for (int q = tid; q < mc * nrhs; q += nt) {
            const int ii = q / nrhs, r = q - ii * nrhs;
            const int pos = ri[ii];
            if (pos < w)
                Y[(int64_t)(c0 + pos) * nrhs + r] -= uc[q];
            else
                us[(pos - w) * nrhs + r] += uc[q];
        }
        __syncthreads();
    }
    if (w > NB) return;  // a wide front: its blocks are the step launches' work
    T* y = Y + (int64_t)c0 * nrhs;
    const bool fuse = m <= FUSE_ROWS;
    const int rend = fuse ? f : w;  // rows this workgroup updates
    for (int k = 0; k < w; k += NB) {
        const int nb = min(NB, w - k);
        for (int q = tid; q < nb * nrhs; q += nt) xs[q] = y[k * nrhs + q];
        __syncthreads();
        block_gemv<T>(DB + (int64_t)(c0 + k) * NB, xs, ys, nb, nrhs, 0);
        __syncthreads();
        for (int q = tid; q < nb * nrhs; q += nt) y[k * nrhs + q] = ys[q];
        // rows below the block: in L11 (subtract now) or below the front (update vector)
        for (int q = tid; q < (rend - k - nb) * nrhs; q += nt) {
            const int i = k + nb + q / nrhs, r = q % nrhs;
            T acc = T(0);
#pragma unroll 8
            for (int j = 0; j < nb; ++j) acc += P[(int64_t)(k + j) * f + i] * ys[j * nrhs + r];
            if (i < w)
                y[i * nrhs + r] -= acc;
            else if (us)
                us[(i - w) * nrhs + r] += acc;
        }
        __syncthreads();
    }
}

// U[rows >= w] += L21 y_top for the tall fronts: a tile of L21_ROWS rows per workgroup, one
// row per thread, all nrhs right-hand sides 8 at a time. task: (s, row0)
template <typename T>
__global__ void __launch_bounds__(L21_ROWS) k_fwd_l21(DevSym S, const Task* __restrict__ tasks,
                                                     const T* __restrict__ factor, const T* __restrict__ Y,
                                                     T* __restrict__ U, int nrhs) {
    __shared__ T yt[NB * MAX_NRHS];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    const int c0 = S.super_ptr[s];
    const int i = t.a + threadIdx.x;
    const T* P = factor + S.panel_off[s];
    T* us = U + S.uv_off[s] * nrhs;
    const T* y = Y + (int64_t)c0 * nrhs;
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
        for (int k = 0; k < w; k += NB) {
            const int nb = min(NB, w - k);
            __syncthreads();
            for (int q = threadIdx.x; q < nb * 8; q += L21_ROWS) {
                const int j = q >> 3, r = q & 7;
                yt[q] = rb + r < nrhs ? y[(k + j) * nrhs + rb + r] : T(0);
            }
            __syncthreads();
            if (i < f) {
#pragma unroll 8
                for (int j = 0; j < nb; ++j) {
                    const T l = P[(int64_t)(k + j) * f + i];
#pragma unroll
                    for (int r = 0; r < 8; ++r) acc[r] += l * yt[j * 8 + r];
                }
            }
        }
        if (i < f)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (rb + r < nrhs) us[(int64_t)(i - w) * nrhs + rb + r] += acc[r];
    }
}

// ---------------------------------------------------------------------------
// backward

// partial sums of L21^T x over a chunk of L21_T_CHUNK rows for a block of NB columns of a tall
// front: part[(chunk * w + j) * nrhs + r]. Thread (column jj, rhs pair) accumulates over the
// rows; L21 tiles and the gathered solution rows are staged in LDS. task: (s, col0, chunk, poff)
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_l21t(DevSym S, const Task* __restrict__ tasks,
                                                 const T* __restrict__ factor, const T* __restrict__ Y,
                                                 T* __restrict__ part, int nrhs) {
    __shared__ T Lt[64][65];
    __shared__ T Yc[64][8];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], col0 = t.a, chunk = t.b;
    const int ncol = min(64, w - col0);
    const int tid = threadIdx.x, jj = tid & 63, rr = (tid >> 6) * 2;
    const T* P = factor + S.panel_off[s];
    const int32_t* rows = S.rowval + S.rowptr[s];
    const int r0 = w + chunk * L21T_CHUNK, r1 = min(f, r0 + L21T_CHUNK);
    T* out = part + (int64_t)t.c * nrhs;  // this front's partials start at t.c (in units of nrhs)
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc0 = T(0), acc1 = T(0);
        for (int i0 = r0; i0 < r1; i0 += 64) {
            const int ni = min(64, r1 - i0);
#pragma unroll
            for (int q = tid; q < 64 * 64; q += 256) {
                const int c = q >> 6, ii = q & 63;
                Lt[c][ii] = (c < ncol && ii < ni) ? P[(int64_t)(col0 + c) * f + i0 + ii] : T(0);
            }
            for (int q = tid; q < 64 * 8; q += 256) {
                const int ii = q >> 3, r = q & 7;
                Yc[ii][r] = (ii < ni && rb + r < nrhs) ? Y[(int64_t)rows[i0 + ii] * nrhs + rb + r] : T(0);
            }
            __syncthreads();
#pragma unroll 8
            for (int ii = 0; ii < 64; ++ii) {
                const T l = Lt[jj][ii];
                acc0 += l * Yc[ii][rr];
                acc1 += l * Yc[ii][rr + 1];
            }
            __syncthreads();
        }
        if (jj < ncol) {
            const int64_t o = ((int64_t)chunk * w + col0 + jj) * nrhs + rb + rr;
            if (rb + rr < nrhs) out[o] = acc0;
            if (rb + rr + 1 < nrhs) out[o + 1] = acc1;
        }
    }
}

// one workgroup per front of the level: y_top -= L21^T x_bottom (from the partials of the tall
// fronts, in chunk order, or computed here), then x_top = L11^-T y_top block by block from the
// last. poff[s]: the front's partial-sum offset (units of nrhs), -1 when computed here.
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_front(DevSym S, const int32_t* __restrict__ fronts,
                                                  const int64_t* __restrict__ poff, const T* __restrict__ factor,
                                                  const T* __restrict__ DB, T* __restrict__ Y,
                                                  const T* __restrict__ part, int nrhs) {
    __shared__ T xs[NB * MAX_NRHS];
    __shared__ T ys[NB * MAX_NRHS];
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const T* P = factor + S.panel_off[s];
    const int32_t* rows = S.rowval + S.rowptr[s];
    T* y = Y + (int64_t)c0 * nrhs;
    const int64_t po = poff[s];
    if (m > 0) {
        if (po >= 0) {
            const int nch = (m + L21T_CHUNK - 1) / L21T_CHUNK;
            const T* pp = part + po * nrhs;
            for (int q = tid; q < w * nrhs; q += nt) {
                T acc = T(0);
                for (int ch = 0; ch < nch; ++ch) acc += pp[(int64_t)ch * w * nrhs + q];
                y[q] -= acc;
            }
        } else {
            // L21^T x over 64-row tiles staged in LDS: thread (column jj, right-hand-side pair)
            // (w <= NB here: a wide front's partial sums are always computed apart)
            __shared__ T Lt[NB][NB + 1];
            __shared__ T Yc[NB][8];
            const int jj = tid & (NB - 1), rr = (tid / NB) * 2;
            for (int rb = 0; rb < nrhs; rb += 8) {
                T acc0 = T(0), acc1 = T(0);
                for (int i0 = w; i0 < f; i0 += NB) {
                    const int ni = min(NB, f - i0);
                    for (int q = tid; q < NB * NB; q += nt) {
                        const int c = q / NB, ii = q % NB;
                        Lt[c][ii] = (c < w && ii < ni) ? P[(int64_t)c * f + i0 + ii] : T(0);
                    }
                    for (int q = tid; q < NB * 8; q += nt) {
                        const int ii = q >> 3, r = q & 7;
                        Yc[ii][r] = (ii < ni && rb + r < nrhs) ? Y[(int64_t)rows[i0 + ii] * nrhs + rb + r] : T(0);
                    }
                    __syncthreads();
#pragma unroll 8
                    for (int ii = 0; ii < NB; ++ii) {
                        const T l = Lt[jj][ii];
                        acc0 += l * Yc[ii][rr];
                        acc1 += l * Yc[ii][rr + 1];
                    }
                    __syncthreads();
                }
                if (jj < w) {
                    if (rb + rr < nrhs) y[jj * nrhs + rb + rr] -= acc0;
                    if (rb + rr + 1 < nrhs) y[jj * nrhs + rb + rr + 1] -= acc1;
                }
            }
        }
        __syncthreads();
    }
    if (w > NB) return;  // a wide front: its blocks are the step launches' work
    const int kl = ((w - 1) / NB) * NB;
    for (int k = kl; k >= 0; k -= NB) {
        const int nb = min(NB, w - k);
        // the block's right-hand side minus L(rows below in L11, block)^T x
        for (int q = tid; q < nb * nrhs; q += nt) {
            const int j = q / nrhs, r = q - j * nrhs;
            T acc = T(0);
            for (int i = k + nb; i < w; ++i) acc += P[(int64_t)(k + j) * f + i] * y[i * nrhs + r];
            xs[q] = y[k * nrhs + q] - acc;
        }
        __syncthreads();
        block_gemv<T>(DB + (int64_t)(c0 + k) * NB, xs, ys, nb, nrhs, 1);
        __syncthreads();
        for (int q = tid; q < nb * nrhs; q += nt) y[k * nrhs + q] = ys[q];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// wide fronts: block steps

// forward block step k of a wide front: y_k = Linv_kk y_k (every workgroup), the task with
// c = 1 keeps it in Z; this workgroup's 256 rows below the block get L(rows, block) y_k:
// rows < w of Y, rows >= w of the front's update vector. task: (s, k, row0, c)
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_wstep(DevSym S, const Task* __restrict__ tasks,
                                                  const T* __restrict__ factor, const T* __restrict__ DB,
                                                  T* __restrict__ Y, T* __restrict__ U, T* __restrict__ Z, int nrhs) {
    __shared__ T xs[NB * MAX_NRHS];
    __shared__ T ys[NB * MAX_NRHS];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a, row0 = t.b;
    const int nb = min(NB, w - k);
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    T* y = Y + (int64_t)c0 * nrhs;
    for (int q = tid; q < nb * nrhs; q += 256) xs[q] = y[k * nrhs + q];
    __syncthreads();
    block_gemv<T>(DB + (int64_t)(c0 + k) * NB, xs, ys, nb, nrhs, 0);
    __syncthreads();
    if (t.c)
        for (int q = tid; q < nb * nrhs; q += 256) Z[(int64_t)(c0 + k) * nrhs + q] = ys[q];
    const int i = row0 + tid;
    if (i >= f) return;
    T* us = S.uv_off[s] >= 0 ? U + S.uv_off[s] * nrhs : nullptr;
    if (i >= w && !us) return;
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
#pragma unroll 8
        for (int j = 0; j < nb; ++j) {
            const T l = P[(int64_t)(k + j) * f + i];
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] += l * ys[j * nrhs + min(rb + r, nrhs - 1)];
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (rb + r >= nrhs) break;
            if (i < w)
                y[(int64_t)i * nrhs + rb + r] -= acc[r];
            else
                us[(int64_t)(i - w) * nrhs + rb + r] += acc[r];
        }
    }
}

// backward block step k of a wide front: x_k = Linv_kk^T y_k (every workgroup), the task with
// c = 1 keeps it in Z; this workgroup's 64 earlier columns [col0, col0 + 64) ∩ [0, k) get
// y_j -= sum_i L(k + i, j) x_i. task: (s, k, col0, c)
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_wstep(DevSym S, const Task* __restrict__ tasks,
                                                  const T* __restrict__ factor, const T* __restrict__ DB,
                                                  T* __restrict__ Y, T* __restrict__ Z, int nrhs) {
    __shared__ T xs[NB * MAX_NRHS];
    __shared__ T ys[NB * MAX_NRHS];
    __shared__ T Lt[NB][NB + 1];  // Lt[jj][i] = L(k + i, col0 + jj)
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a, col0 = t.b;
    const int nb = min(NB, w - k);
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    T* y = Y + (int64_t)c0 * nrhs;
    const int ncol = max(0, min(col0 + NB, k) - col0);
    for (int q = tid; q < nb * nrhs; q += 256) xs[q] = y[k * nrhs + q];
    for (int q = tid; q < NB * NB; q += 256) {
        const int c = q / NB, i = q % NB;
        Lt[c][i] = (c < ncol && i < nb) ? P[(int64_t)(col0 + c) * f + k + i] : T(0);
    }
    __syncthreads();
    block_gemv<T>(DB + (int64_t)(c0 + k) * NB, xs, ys, nb, nrhs, 1);
    __syncthreads();
    if (t.c)
        for (int q = tid; q < nb * nrhs; q += 256) Z[(int64_t)(c0 + k) * nrhs + q] = ys[q];
    for (int q = tid; q < ncol * nrhs; q += 256) {
        const int jj = q / nrhs, r = q - jj * nrhs;
        T acc = T(0);
#pragma unroll 8
        for (int i = 0; i < nb; ++i) acc += Lt[jj][i] * ys[i * nrhs + r];
        y[(int64_t)(col0 + jj) * nrhs + r] -= acc;
    }
}

// ---------------------------------------------------------------------------

template <typename T>
void launch_block_inverse(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* DB) {
    hipLaunchKernelGGL(k_block_inverse<T>, dim3(n), dim3(64), 0, st, S, tasks, factor, DB);
}

template <typename T>
void launch_fwd_front(hipStream_t st, const DevSym& S, const int32_t* fronts, int n, const T* factor, const T* DB, T* Y,
                      T* U, int nrhs) {
    hipLaunchKernelGGL(k_fwd_front<T>, dim3(n), dim3(256), 0, st, S, fronts, factor, DB, Y, U, nrhs);
}

template <typename T>
void launch_fwd_l21(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* Y, T* U,
                    int nrhs) {
    hipLaunchKernelGGL(k_fwd_l21<T>, dim3(n), dim3(L21_ROWS), 0, st, S, tasks, factor, Y, U, nrhs);
}

template <typename T>
void launch_bwd_l21t(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* Y, T* part,
                     int nrhs) {
    hipLaunchKernelGGL(k_bwd_l21t<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, Y, part, nrhs);
}

template <typename T>
void launch_fwd_wstep(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* DB, T* Y,
                      T* U, T* Z, int nrhs) {
    hipLaunchKernelGGL(k_fwd_wstep<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, DB, Y, U, Z, nrhs);
}

template <typename T>
void launch_bwd_wstep(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* DB, T* Y,
                      T* Z, int nrhs) {
    hipLaunchKernelGGL(k_bwd_wstep<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, DB, Y, Z, nrhs);
}

template <typename T>
void launch_bwd_front(hipStream_t st, const DevSym& S, const int32_t* fronts, int n, const int64_t* poff,
                      const T* factor, const T* DB, T* Y, const T* part, int nrhs) {
    hipLaunchKernelGGL(k_bwd_front<T>, dim3(n), dim3(256), 0, st, S, fronts, poff, factor, DB, Y, part, nrhs);
}

#define INST(T)                                                                                                     \
    template void launch_block_inverse<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*);             \
    template void launch_fwd_front<T>(hipStream_t, const DevSym&, const int32_t*, int, const T*, const T*, T*, T*, \
                                      int);                                                                        \
    template void launch_fwd_l21<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, int);    \
    template void launch_bwd_l21t<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, int);   \
    template void launch_bwd_front<T>(hipStream_t, const DevSym&, const int32_t*, int, const int64_t*, const T*,   \
                                      const T*, T*, const T*, int);                                                \
    template void launch_fwd_wstep<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, T*, T*, \
                                      int);                                                                        \
    template void launch_bwd_wstep<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, T*, int);
INST(double)
INST(float)
#undef INST

}  // namespace sds
