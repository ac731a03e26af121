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
// Per tree level (narrow fronts: w <= NB columns, one diagonal block; wide fronts: w > NB):
//   forward : k_fwd_front, one workgroup per narrow front: pull the children's update
//             vectors, y_top = Linv y_top, and, when the front has at most FUSE_ROWS rows below
//             its columns, its update vector u = L21 y_top; k_fwd_sf, one wave per diagonal
//             block of the wide fronts: block k waits for the blocks before it (flags) and forms
//             y_k = Linv_kk (y_k - sum_j L_kj y_j); k_fwd_l21, one wave per 64 rows below the
//             columns of the remaining fronts: pull the children's entries of those rows and add
//             L21 y_top.
//   backward: k_bwd_l21t, partial sums of L21^T x per chunk of L21T_CHUNK rows (one wave per
//             group of columns, lanes along the rows, a reduce-scatter across the lanes at the
//             end); k_bwd_front, one workgroup per narrow front: y_top -= L21^T x (partials in
//             chunk order, or LDS-staged rows), x_top = Linv^T y_top; k_bwd_sf, one wave per
//             diagonal block of the wide fronts from the last: x_k = Linv_kk^T (y_k - partials
//             - sum_{j>k} L_jk^T x_j).
// The flags of the wide fronts' blocks are cleared at the start of each sweep; a block waits
// only for blocks with lower workgroup index in the same launch, and workgroups are dispatched
// in index order, so every wait ends (the spin is bounded all the same: on a timeout the wave
// records an error word and goes on).
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

extern __shared__ __align__(16) unsigned char sweep_lds[];

// first index in the ascending ri[0, n) with ri[idx] >= v
__device__ inline int lower_bound_i32(const int32_t* __restrict__ ri, int n, int v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ri[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

constexpr int SPIN_LIMIT = 1 << 22;

// wait (all lanes of the wave) until flag != 0 with agent-scope acquire ordering
__device__ inline void wait_flag(const int32_t* flag, int32_t* err) {
    int spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++spins > SPIN_LIMIT) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// publish the block's results (all lanes have stored them) with agent-scope release ordering
__device__ inline void set_flag(int32_t* flag) {
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

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

// ---------------------------------------------------------------------------
// forward

// one workgroup per narrow front of the level (w <= NB). Fused fronts (m <= FUSE_ROWS) also
// form their update vector; the others leave it to k_fwd_l21 (rows >= w, pulls included).
// LDS: 2 NB nrhs entries.
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_front(DevSym S, const int32_t* __restrict__ fronts,
                                                  const T* __restrict__ factor, const T* __restrict__ DB,
                                                  T* __restrict__ Y, T* __restrict__ U, int nrhs) {
    T* xs = reinterpret_cast<T*>(sweep_lds);
    T* ys = xs + NB * nrhs;
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const T* P = factor + S.panel_off[s];
    const bool fused = m <= FUSE_ROWS;
    T* us = fused && S.uv_off[s] >= 0 ? U + S.uv_off[s] * nrhs : nullptr;
    if (us)
        for (int q = tid; q < m * nrhs; q += nt) us[q] = T(0);
    __syncthreads();
    // the children's update vectors, in child_list order (only the rows < w when not fused)
    for (int kc = S.child_ptr[s]; kc < S.child_ptr[s + 1]; ++kc) {
        const int c = S.child_list[kc];
        if (S.uv_off[c] < 0) continue;
        const int mc = S.fnrows[c] - S.fncols[c];
        const T* uc = U + S.uv_off[c] * nrhs;
        const int32_t* ri = S.relind + S.relind_ptr[c];
        const int hi = us ? mc : lower_bound_i32(ri, mc, w);
        for (int q = tid; q < hi * nrhs; q += nt) {
            const int ii = q / nrhs, r = q - ii * nrhs;
            const int pos = ri[ii];
            if (pos < w)
                Y[(int64_t)(c0 + pos) * nrhs + r] -= uc[q];
            else
                us[(pos - w) * nrhs + r] += uc[q];
        }
        __syncthreads();
    }
    T* y = Y + (int64_t)c0 * nrhs;
    for (int q = tid; q < w * nrhs; q += nt) xs[q] = y[q];
    __syncthreads();
    block_gemv<T>(DB + (int64_t)c0 * NB, xs, ys, w, nrhs, 0);
    __syncthreads();
    for (int q = tid; q < w * nrhs; q += nt) y[q] = ys[q];
    if (!us) return;
    // update vector rows, one per thread, 8 right-hand sides per pass over the row's entries
    for (int i = w + tid; i < f; i += nt)
        for (int rb = 0; rb < nrhs; rb += 8) {
            T acc[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] = T(0);
#pragma unroll 8
            for (int j = 0; j < w; ++j) {
                const T l = P[(int64_t)j * f + i];
#pragma unroll
                for (int r = 0; r < 8; ++r) acc[r] += l * ys[j * nrhs + min(rb + r, nrhs - 1)];
            }
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (rb + r < nrhs) us[(i - w) * nrhs + rb + r] += acc[r];
        }
}

// one wave per diagonal block of the wide fronts: task (s, k, flag base, 0). The block's rows
// get their children's entries, then y_k -= L_kj y_j for every earlier block j (after its
// flag), then y_k = Linv_kk y_k, and the block's flag is set. LDS: NB nrhs + NB 8 entries.
template <typename T>
__global__ void __launch_bounds__(64) k_fwd_sf(DevSym S, const Task* __restrict__ tasks, const T* __restrict__ factor,
                                              const T* __restrict__ DB, T* __restrict__ Y, const T* __restrict__ U,
                                              int32_t* __restrict__ flags, int32_t* __restrict__ err, int nrhs) {
    T* ts = reinterpret_cast<T*>(sweep_lds);  // NB x nrhs
    T* yj = ts + NB * nrhs;                   // NB x 8
    const Task t = tasks[blockIdx.x];
    const int s = t.s, k = t.a;
    const int f = S.fnrows[s], w = S.fncols[s];
    const int c0 = S.super_ptr[s];
    const int nb = min(NB, w - k);
    const int lane = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    T* y = Y + (int64_t)c0 * nrhs;
    for (int kc = S.child_ptr[s]; kc < S.child_ptr[s + 1]; ++kc) {
        const int c = S.child_list[kc];
        if (S.uv_off[c] < 0) continue;
        const int mc = S.fnrows[c] - S.fncols[c];
        const T* uc = U + S.uv_off[c] * nrhs;
        const int32_t* ri = S.relind + S.relind_ptr[c];
        const int lo = lower_bound_i32(ri, mc, k), hi = lower_bound_i32(ri, mc, k + nb);
        for (int q = lane; q < (hi - lo) * nrhs; q += 64) {
            const int ii = lo + q / nrhs, r = q % nrhs;
            y[(int64_t)ri[ii] * nrhs + r] -= uc[(int64_t)ii * nrhs + r];
        }
        __syncthreads();
    }
    const int i = k + min(lane, nb - 1);  // this lane's row (clamped; results of lanes >= nb unused)
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
        for (int j = 0; j < k; j += NB) {
            T l[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) l[c] = P[(int64_t)(j + c) * f + i];
            wait_flag(flags + t.b + j / NB, err);
#pragma unroll
            for (int r = 0; r < 8; ++r) yj[lane * 8 + r] = rb + r < nrhs ? y[(int64_t)(j + lane) * nrhs + rb + r] : T(0);
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NB; ++c)
#pragma unroll
                for (int r = 0; r < 8; ++r) acc[r] += l[c] * yj[c * 8 + r];
            __syncthreads();
        }
        if (lane < nb)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (rb + r < nrhs) ts[lane * nrhs + rb + r] = y[(int64_t)i * nrhs + rb + r] - acc[r];
    }
    __syncthreads();
    const T* Li = DB + (int64_t)(c0 + k) * NB;
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
        for (int ii = 0; ii < nb; ++ii) {
            const T li = Li[(int64_t)ii * NB + lane];  // Linv(lane, ii): zero above the diagonal
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] += li * ts[ii * nrhs + min(rb + r, nrhs - 1)];
        }
        if (lane < nb)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (rb + r < nrhs) y[(int64_t)(k + lane) * nrhs + rb + r] = acc[r];
    }
    set_flag(flags + t.b + k / NB);
}

// the update vector rows [r0, r0 + 64) (r0 >= w) of a front that is wide or tall, one wave:
// zero them, add the children's entries of those rows (child_list order), add L21 y_top.
// task: (s, r0)
template <typename T>
__global__ void __launch_bounds__(64) k_fwd_l21(DevSym S, const Task* __restrict__ tasks,
                                               const T* __restrict__ factor, const T* __restrict__ Y,
                                               T* __restrict__ U, int nrhs) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    const int c0 = S.super_ptr[s];
    const int r0 = t.a, r1 = min(f, r0 + 64);
    const int lane = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    T* us = U + S.uv_off[s] * nrhs;
    for (int q = lane; q < (r1 - r0) * nrhs; q += 64) us[(int64_t)(r0 - w) * nrhs + q] = T(0);
    __syncthreads();
    for (int kc = S.child_ptr[s]; kc < S.child_ptr[s + 1]; ++kc) {
        const int c = S.child_list[kc];
        if (S.uv_off[c] < 0) continue;
        const int mc = S.fnrows[c] - S.fncols[c];
        const T* uc = U + S.uv_off[c] * nrhs;
        const int32_t* ri = S.relind + S.relind_ptr[c];
        const int lo = lower_bound_i32(ri, mc, r0), hi = lower_bound_i32(ri, mc, r1);
        for (int q = lane; q < (hi - lo) * nrhs; q += 64) {
            const int ii = lo + q / nrhs, r = q % nrhs;
            us[(int64_t)(ri[ii] - w) * nrhs + r] += uc[(int64_t)ii * nrhs + r];
        }
        __syncthreads();
    }
    const int i = min(r0 + lane, f - 1);
    const T* y = Y + (int64_t)c0 * nrhs;
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
#pragma unroll 16
        for (int j = 0; j < w; ++j) {
            const T l = P[(int64_t)j * f + i];
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] += l * y[(int64_t)j * nrhs + min(rb + r, nrhs - 1)];
        }
        if (r0 + lane < f)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (rb + r < nrhs) us[(int64_t)(i - w) * nrhs + rb + r] += acc[r];
    }
}

// ---------------------------------------------------------------------------
// backward

// partial sums of L21^T x over the chunk of L21T_CHUNK rows of a task (s, col0, chunk): a wave
// per CW = 64 / RW columns; lanes along the rows; RW right-hand sides per pass. The 64 sums
// (CW columns x RW right-hand sides) are reduce-scattered across the lanes (recursive
// halving: lane l ends with sum l). part[(poff + chunk w + j) nrhs + r].
template <typename T, int RW>
__global__ void __launch_bounds__(64) k_bwd_l21t(DevSym S, const Task* __restrict__ tasks,
                                                const int64_t* __restrict__ poff, const T* __restrict__ factor,
                                                const T* __restrict__ Y, T* __restrict__ part, int nrhs) {
    constexpr int CW = 64 / RW;
    const Task t = tasks[blockIdx.x / RW];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], chunk = t.b;
    const int col0 = t.a + (blockIdx.x % RW) * CW;
    if (col0 >= w) return;
    const int lane = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    const int32_t* rows = S.rowval + S.rowptr[s];
    const int r0 = w + chunk * L21T_CHUNK, r1 = min(f, r0 + L21T_CHUNK);
    T* out = part + (poff[s] + (int64_t)chunk * w) * nrhs;
    for (int rb = 0; rb < nrhs; rb += RW) {
        T a[64];
#pragma unroll
        for (int v = 0; v < 64; ++v) a[v] = T(0);
#pragma unroll(RW >= 4 ? 2 : 1)
        for (int i = r0 + lane; i < r1; i += 64) {
            const T* yr = Y + (int64_t)rows[i] * nrhs;
            T x[RW];
#pragma unroll
            for (int r = 0; r < RW; ++r) x[r] = yr[min(rb + r, nrhs - 1)];
            const T* pc = P + (int64_t)col0 * f + i;  // column col0 + c; columns past w read w - 1
#pragma unroll
            for (int c = 0; c < CW; ++c) {
                const T l = *pc;
                pc += col0 + c + 1 < w ? f : 0;
#pragma unroll
                for (int r = 0; r < RW; ++r) a[c * RW + r] += l * x[r];
            }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const bool up = (lane & d) != 0;
#pragma unroll
            for (int v = 0; v < d; ++v) {
                const T send = up ? a[v] : a[v + d];
                const T keep = up ? a[v + d] : a[v];
                a[v] = keep + __shfl_xor(send, d);
            }
        }
        const int c = lane / RW, r = lane % RW;
        if (col0 + c < w && rb + r < nrhs) out[(int64_t)(col0 + c) * nrhs + rb + r] = a[0];
    }
}

// one workgroup per narrow front of the level: y_top -= L21^T x (partials in chunk order, or
// from LDS-staged solution rows when fused), then x_top = Linv^T y_top.
// LDS: 3 NB nrhs entries.
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_front(DevSym S, const int32_t* __restrict__ fronts,
                                                  const int64_t* __restrict__ poff, const T* __restrict__ factor,
                                                  const T* __restrict__ DB, T* __restrict__ Y,
                                                  const T* __restrict__ part, int nrhs) {
    T* ts = reinterpret_cast<T*>(sweep_lds);
    T* ys = ts + NB * nrhs;
    T* Xs = ys + NB * nrhs;
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const T* P = factor + S.panel_off[s];
    const int32_t* rows = S.rowval + S.rowptr[s];
    T* y = Y + (int64_t)c0 * nrhs;
    for (int q = tid; q < w * nrhs; q += nt) ts[q] = y[q];
    __syncthreads();
    if (m > 0) {
        const int64_t po = poff[s];
        if (po >= 0) {
            const int nch = (m + L21T_CHUNK - 1) / L21T_CHUNK;
            const T* pp = part + po * nrhs;
            for (int q = tid; q < w * nrhs; q += nt) {
                T v = ts[q];
                for (int ch = 0; ch < nch; ++ch) v -= pp[(int64_t)ch * w * nrhs + q];
                ts[q] = v;
            }
        } else {
            for (int i0 = w; i0 < f; i0 += 64) {
                const int ni = min(64, f - i0);
                for (int q = tid; q < ni * nrhs; q += nt) {
                    const int ii = q / nrhs, r = q - ii * nrhs;
                    Xs[q] = Y[(int64_t)rows[i0 + ii] * nrhs + r];
                }
                __syncthreads();
                for (int q = tid; q < w * nrhs; q += nt) {
                    const int j = q / nrhs, r = q - j * nrhs;
                    const T* pj = P + (int64_t)j * f + i0;
                    T acc = T(0);
                    for (int ii = 0; ii < ni; ++ii) acc += pj[ii] * Xs[ii * nrhs + r];
                    ts[q] -= acc;
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    block_gemv<T>(DB + (int64_t)c0 * NB, ts, ys, w, nrhs, 1);
    __syncthreads();
    for (int q = tid; q < w * nrhs; q += nt) y[q] = ys[q];
}

// one wave per diagonal block of the wide fronts, the blocks of a front from the last: task
// (s, k, flag base, 0). t = y_k - L21^T x partials (chunk order) - sum_{j > k, descending}
// L_jk^T x_j (after block j's flag; L_jk staged through LDS), x_k = Linv_kk^T t, flag set.
// LDS: NB nrhs + NB (NB + 1) + NB 8 entries.
template <typename T>
__global__ void __launch_bounds__(64) k_bwd_sf(DevSym S, const Task* __restrict__ tasks,
                                              const int64_t* __restrict__ poff, const T* __restrict__ factor,
                                              const T* __restrict__ DB, T* __restrict__ Y, const T* __restrict__ part,
                                              int32_t* __restrict__ flags, int32_t* __restrict__ err, int nrhs) {
    T* ts = reinterpret_cast<T*>(sweep_lds);  // NB x nrhs
    T(*Lt)[NB + 1] = reinterpret_cast<T(*)[NB + 1]>(ts + NB * nrhs);
    T* xj = ts + NB * nrhs + NB * (NB + 1);  // NB x 8
    const Task t = tasks[blockIdx.x];
    const int s = t.s, k = t.a;
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int nb = min(NB, w - k);
    const int lane = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    T* y = Y + (int64_t)c0 * nrhs;
    const int64_t po = poff[s];
    const int nch = m > 0 && po >= 0 ? (m + L21T_CHUNK - 1) / L21T_CHUNK : 0;
    for (int q = lane; q < nb * nrhs; q += 64) {
        T v = y[(int64_t)k * nrhs + q];
        for (int ch = 0; ch < nch; ++ch) v -= part[(po + (int64_t)ch * w + k) * nrhs + q];
        ts[q] = v;
    }
    const int cl = min(lane, nb - 1);  // this lane's column (clamped; lanes >= nb unused)
    const int kl = ((w - 1) / NB) * NB;
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
        for (int j = kl; j > k; j -= NB) {
            const int njb = min(NB, w - j);
            const int ir = j + min(lane, njb - 1);
            __syncthreads();
#pragma unroll 16
            for (int c = 0; c < NB; ++c) Lt[c][lane] = lane < njb ? P[(int64_t)(k + min(c, nb - 1)) * f + ir] : T(0);
            wait_flag(flags + t.b + j / NB, err);
#pragma unroll
            for (int r = 0; r < 8; ++r) xj[lane * 8 + r] = rb + r < nrhs ? y[(int64_t)ir * nrhs + rb + r] : T(0);
            __syncthreads();
#pragma unroll 8
            for (int ii = 0; ii < NB; ++ii) {
                const T l = Lt[cl][ii];
#pragma unroll
                for (int r = 0; r < 8; ++r) acc[r] += l * xj[ii * 8 + r];
            }
        }
        __syncthreads();
        if (lane < nb)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (rb + r < nrhs) ts[lane * nrhs + rb + r] -= acc[r];
    }
    // Linv^T: Lt[c][i] = Linv(i, c)
    const T* Li = DB + (int64_t)(c0 + k) * NB;
    __syncthreads();
#pragma unroll 16
    for (int c = 0; c < NB; ++c) Lt[c][lane] = c < nb ? Li[(int64_t)c * NB + lane] : T(0);
    __syncthreads();
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
        for (int ii = 0; ii < nb; ++ii) {
            const T l = Lt[cl][ii];  // zero above the diagonal
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] += l * ts[ii * nrhs + min(rb + r, nrhs - 1)];
        }
        if (lane < nb)
#pragma unroll
            for (int r = 0; r < 8; ++r)
                if (rb + r < nrhs) y[(int64_t)(k + lane) * nrhs + rb + r] = acc[r];
    }
    set_flag(flags + t.b + k / NB);
}

// ---------------------------------------------------------------------------

// dynamic LDS beyond 64 KiB has to be allowed per kernel (once per process and kernel)
template <typename K>
void allow_lds(K kernel, size_t bytes) {
    if (bytes > 65536) (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename T>
void launch_block_inverse(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* DB) {
    hipLaunchKernelGGL(k_block_inverse<T>, dim3(n), dim3(64), 0, st, S, tasks, factor, DB);
}

template <typename T>
void launch_fwd_front(hipStream_t st, const DevSym& S, const int32_t* fronts, int n, int threads, const T* factor,
                      const T* DB, T* Y, T* U, int nrhs) {
    const size_t lds = 2 * NB * nrhs * sizeof(T);
    allow_lds(k_fwd_front<T>, lds);
    hipLaunchKernelGGL(k_fwd_front<T>, dim3(n), dim3(threads), lds, st, S, fronts, factor, DB, Y, U, nrhs);
}

template <typename T>
void launch_fwd_sf(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* DB, T* Y,
                   const T* U, int32_t* flags, int32_t* err, int nrhs) {
    hipLaunchKernelGGL(k_fwd_sf<T>, dim3(n), dim3(64), (NB * nrhs + NB * 8) * sizeof(T), st, S, tasks, factor, DB, Y,
                       U, flags, err, nrhs);
}

template <typename T>
void launch_fwd_l21(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* Y, T* U,
                    int nrhs) {
    hipLaunchKernelGGL(k_fwd_l21<T>, dim3(n), dim3(64), 0, st, S, tasks, factor, Y, U, nrhs);
}

template <typename T>
void launch_bwd_l21t(hipStream_t st, const DevSym& S, const Task* tasks, int n, const int64_t* poff, const T* factor,
                     const T* Y, T* part, int nrhs) {
    if (nrhs >= 8)
        hipLaunchKernelGGL((k_bwd_l21t<T, 8>), dim3(n * 8), dim3(64), 0, st, S, tasks, poff, factor, Y, part, nrhs);
    else if (nrhs >= 4)
        hipLaunchKernelGGL((k_bwd_l21t<T, 4>), dim3(n * 4), dim3(64), 0, st, S, tasks, poff, factor, Y, part, nrhs);
    else if (nrhs >= 2)
        hipLaunchKernelGGL((k_bwd_l21t<T, 2>), dim3(n * 2), dim3(64), 0, st, S, tasks, poff, factor, Y, part, nrhs);
    else
        hipLaunchKernelGGL((k_bwd_l21t<T, 1>), dim3(n), dim3(64), 0, st, S, tasks, poff, factor, Y, part, nrhs);
}

template <typename T>
void launch_bwd_front(hipStream_t st, const DevSym& S, const int32_t* fronts, int n, int threads, const int64_t* poff,
                      const T* factor, const T* DB, T* Y, const T* part, int nrhs) {
    const size_t lds = 3 * NB * nrhs * sizeof(T);
    allow_lds(k_bwd_front<T>, lds);
    hipLaunchKernelGGL(k_bwd_front<T>, dim3(n), dim3(threads), lds, st, S, fronts, poff, factor, DB, Y, part, nrhs);
}

template <typename T>
void launch_bwd_sf(hipStream_t st, const DevSym& S, const Task* tasks, int n, const int64_t* poff, const T* factor,
                   const T* DB, T* Y, const T* part, int32_t* flags, int32_t* err, int nrhs) {
    const size_t lds = (NB * nrhs + NB * (NB + 1) + NB * 8) * sizeof(T);
    allow_lds(k_bwd_sf<T>, lds);
    hipLaunchKernelGGL(k_bwd_sf<T>, dim3(n), dim3(64), lds, st, S, tasks, poff, factor, DB, Y, part, flags, err, nrhs);
}

#define INST(T)                                                                                                     \
    template void launch_block_inverse<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*);                   \
    template void launch_fwd_front<T>(hipStream_t, const DevSym&, const int32_t*, int, int, const T*, const T*, T*, \
                                      T*, int);                                                                    \
    template void launch_fwd_sf<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, const T*,  \
                                   int32_t*, int32_t*, int);                                                       \
    template void launch_fwd_l21<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, int);    \
    template void launch_bwd_l21t<T>(hipStream_t, const DevSym&, const Task*, int, const int64_t*, const T*,       \
                                     const T*, T*, int);                                                           \
    template void launch_bwd_front<T>(hipStream_t, const DevSym&, const int32_t*, int, int, const int64_t*,        \
                                      const T*, const T*, T*, const T*, int);                                      \
    template void launch_bwd_sf<T>(hipStream_t, const DevSym&, const Task*, int, const int64_t*, const T*,         \
                                   const T*, T*, const T*, int32_t*, int32_t*, int);
INST(double)
INST(float)
#undef INST

}  // namespace sds
