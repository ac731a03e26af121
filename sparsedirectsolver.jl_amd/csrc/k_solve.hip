// Solve-phase kernels (gfx950): right-hand-side permutation, level-batched
// multifrontal forward and backward sweeps, the D^-1 step of LDL^T, and the
// FP64 residual of iterative refinement.
//
// Work vectors are row-interleaved (Y[k * nrhs + r]), so a gather of one row
// reads all its right-hand sides from one cache line. The forward sweep is the
// deterministic owner-pull form: a front pulls its children's update vectors
// (m entries per right-hand side, "to subtract") in child_list order, solves
// with L11 and leaves L21 x (plus what it pulled for its rows below) as its own
// update vector; no atomics. The backward sweep gathers the final values of a
// front's rows below from Y.
//
// Reference counterparts: src/solve/permute.jl:68-90, src/solve/sweeps.jl:438-570,
// src/solve/refinement.jl:136-200.
#include "sds_common.h"
#include "sds_kernels.h"

namespace sds {

// Y[k nrhs + r] = sc_i B[i + r ldb] with i = cq[perm[k]] (cq: the matching's column permutation
// of "G" on the side that needs it, else none; sc: the matching's scaling of that side, or none)
// (src/solve/permute.jl:68-150 with the SolveScaling of src/solver.jl:583-604)
template <typename T>
__global__ void k_permute_in(int n, int nrhs, const int32_t* __restrict__ perm, const int32_t* __restrict__ cq,
                             const double* __restrict__ sc, const double* __restrict__ B, int64_t ldb,
                             T* __restrict__ Y) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / nrhs), r = (int)(q - (int64_t)k * nrhs);
        int i = perm ? perm[k] : k;
        if (cq) i = cq[i];
        const double b = B[i + r * ldb];
        Y[q] = T(sc ? sc[i] * b : b);
    }
}

// the values the factorization reads with a matching: w_p a_p (w_p = r_i c_j of the stored entry)
__global__ void k_scale_values(int64_t nnz, const double* __restrict__ w, const double* __restrict__ nz,
                               double* __restrict__ out) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = w[q] * nz[q];
}

// Schur complement mode, forward sweep: the Schur root is not solved; its rows of Y get their
// children's update vectors (child_list order), leaving the condensed right-hand side
// Y2 - L21 Y1 (src/solve/sweeps.jl:919-925). One workgroup; each row has one writer per child.
template <typename T>
__global__ void __launch_bounds__(256) k_schur_pull(DevSym S, int s, T* __restrict__ Y, const T* __restrict__ U,
                                                   int nrhs) {
    const int cg = S.super_ptr[s];
    for (int k = S.child_ptr[s]; k < S.child_ptr[s + 1]; ++k) {
        const int c = S.child_list[k];
        if (S.uv_off[c] < 0) continue;
        const int mc = S.fnrows[c] - S.fncols[c];
        const T* uc = U + S.uv_off[c] * nrhs;
        const int32_t* ri = S.relind + S.relind_ptr[c];
        for (int q = threadIdx.x; q < mc * nrhs; q += 256) {
            const int ii = q / nrhs, r = q - ii * nrhs;
            Y[(int64_t)(cg + ri[ii]) * nrhs + r] -= uc[q];
        }
        __syncthreads();
    }
}

// X[perm[k], r] = Y[k, r] (add = 0) or X[perm[k], r] += Y[k, r] (add = 1)
// X[i + r ldx] (+)= sc_i Y[k nrhs + r] with i = cq[perm[k]] (as k_permute_in)
template <typename T>
__global__ void k_permute_out(int n, int nrhs, const int32_t* __restrict__ perm, const int32_t* __restrict__ cq,
                              const double* __restrict__ sc, const T* __restrict__ Y, double* __restrict__ X,
                              int64_t ldx, int add) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / nrhs), r = (int)(q - (int64_t)k * nrhs);
        int i = perm ? perm[k] : k;
        if (cq) i = cq[i];
        double* x = X + i + r * ldx;
        const double y = sc ? sc[i] * (double)Y[q] : (double)Y[q];
        *x = add ? *x + y : y;
    }
}

// stage the ni x nj block of panel P (leading dimension f) at (row i0, column j0) into LDS
// as B[jj][ii]; the loop has a compile-time trip count and is unrolled so that each thread
// has all its loads in flight at once (blockDim.x = 256)
template <typename T>
__device__ inline void load_block(T (*B)[NB + 1], const T* __restrict__ P, int f, int i0, int j0, int ni, int nj) {
    // unconditional loads from clamped addresses, then the masked LDS stores: a guarded load
    // made the compiler wait for each load before issuing the next
    T v[NB * NB / 256];
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
        const int q = threadIdx.x + 256 * u, jj = q / NB, ii = q - jj * NB;
        v[u] = P[(int64_t)(j0 + min(jj, max(nj - 1, 0))) * f + i0 + min(ii, max(ni - 1, 0))];  // (nj may be 0)
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
        const int q = threadIdx.x + 256 * u, jj = q / NB, ii = q - jj * NB;
        if (ii < ni && jj < nj) B[jj][ii] = v[u];
    }
}

// solve a diagonal block held in LDS: xs[i * nrhs + r] <- L_bb^-1 xs (lower = 1) or
// L_bb^-T xs (lower = 0); Lb[j][i] = L(k + i, k + j). One wave per right-hand side group,
// lane i holds row i in a register; the solved value moves between lanes by shuffles.
// xs is read before and written after the solve only (callers put barriers around it).
template <typename T>
__device__ void block_solve(const T (*Lb)[NB + 1], T* xs, int nb, int nrhs, int unit, int lower) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int r = wv; r < nrhs; r += nw) {
        T v = lane < nb ? xs[lane * nrhs + r] : T(0);
        if (lower) {
            for (int j = 0; j < nb; ++j) {
                T xj = __shfl(v, j);
                if (!unit) xj /= Lb[j][j];
                if (lane == j) v = xj;
                if (lane > j && lane < nb) v -= Lb[j][lane] * xj;
            }
        } else {
            for (int j = nb - 1; j >= 0; --j) {
                T xj = __shfl(v, j);
                if (!unit) xj /= Lb[j][j];
                if (lane == j) v = xj;
                if (lane < j) v -= Lb[lane][j] * xj;
            }
        }
        if (lane < nb) xs[lane * nrhs + r] = v;
    }
}

// forward sweep over the fronts of one level, one workgroup per front.
// unit: unit diagonal (LDL^T, LU); piv (nullable): local pivot order of each
// supernode (global factor columns), applied to its slice first through the
// scratch Z (same shape as Y; a front only touches its own columns there).
template <typename T>
__global__ void __launch_bounds__(256) k_fwd(DevSym S, const int32_t* __restrict__ fronts,
                                            const T* __restrict__ factor, T* __restrict__ Y, T* __restrict__ U,
                                            T* __restrict__ Z, int nrhs, int unit, const int32_t* __restrict__ piv) {
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    const T* P = factor + S.panel_off[s];
    T* us = S.uv_off[s] >= 0 ? U + S.uv_off[s] * nrhs : nullptr;
    if (us)
        for (int q = tid; q < m * nrhs; q += nt) us[q] = T(0);
    __syncthreads();
    for (int k = S.child_ptr[s]; k < S.child_ptr[s + 1]; ++k) {
        const int c = S.child_list[k];
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
    T* y = Y + (int64_t)c0 * nrhs;
    if (piv) {
        T* z = Z + (int64_t)c0 * nrhs;
        for (int q = tid; q < w * nrhs; q += nt) {
            const int kk = q / nrhs, r = q - kk * nrhs;
            z[q] = Y[(int64_t)piv[c0 + kk] * nrhs + r];
        }
        __syncthreads();
        for (int q = tid; q < w * nrhs; q += nt) y[q] = z[q];
        __syncthreads();
    }
    __shared__ T Lb[NB][NB + 1];
    __shared__ T xs[NB * MAX_NRHS];
    for (int j0 = 0; j0 < w; j0 += NB) {
        const int nb = min(NB, w - j0), j1 = j0 + nb;
        load_block<T>(Lb, P, f, j0, j0, nb, nb);
        for (int q = tid; q < nb * nrhs; q += nt) xs[q] = y[j0 * nrhs + q];
        __syncthreads();
        block_solve<T>(Lb, xs, nb, nrhs, unit, 1);
        __syncthreads();
        for (int q = tid; q < nb * nrhs; q += nt) y[j0 * nrhs + q] = xs[q];
        // rows below the chunk: in L11 (subtract now) or below the supernode (update vector)
        for (int q = tid; q < (f - j1) * nrhs; q += nt) {
            const int i = j1 + q / nrhs, r = q % nrhs;
            T t = T(0);
#pragma unroll 16
            for (int j = 0; j < nb; ++j) t += P[(int64_t)(j0 + j) * f + i] * xs[j * nrhs + r];
            if (i < w)
                y[i * nrhs + r] -= t;
            else if (us)
                us[(i - w) * nrhs + r] += t;
        }
        __syncthreads();
    }
}

// backward sweep over the fronts of one level, one workgroup (4 waves) per front:
// x_j = (y_j - sum_{i > j} L[i, j] x_i) / L[j, j], columns in chunks of 64 from the
// last; the dot products over the rows past the chunk one wave per column, the
// substitution inside the chunk on LDS copies. piv (nullable): undo the local pivot
// order afterwards through the scratch Z.
template <typename T, int RB>
__global__ void __launch_bounds__(256) k_bwd(DevSym S, const int32_t* __restrict__ fronts,
                                            const T* __restrict__ factor, T* __restrict__ Y, T* __restrict__ Z,
                                            int nrhs, int unit, const int32_t* __restrict__ piv) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* xs = reinterpret_cast<T*>(smem_raw);  // 64 * nrhs: the chunk's values
    T* ts = xs + 64 * nrhs;                  // 64 * nrhs: dot products past the chunk
    __shared__ T Lb[NB][NB + 1];             // the chunk's diagonal block
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s];
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const T* P = factor + S.panel_off[s];
    const int32_t* rows = S.rowval + S.rowptr[s];
    T* y = Y + (int64_t)c0 * nrhs;
    for (int j1 = w; j1 > 0; j1 -= 64) {
        const int j0 = max(j1 - 64, 0);
        for (int j = j0 + wv; j < j1; j += 4) {
            for (int rb = 0; rb < nrhs; rb += RB) {
                T acc[RB];
#pragma unroll
                for (int r = 0; r < RB; ++r) acc[r] = T(0);
                for (int i = j1 + lane; i < f; i += 64) {
                    const T l = P[(int64_t)j * f + i];
                    const T* yr = Y + (int64_t)rows[i] * nrhs + rb;
#pragma unroll
                    for (int r = 0; r < RB; ++r)
                        if (rb + r < nrhs) acc[r] += l * yr[r];
                }
#pragma unroll
                for (int r = 0; r < RB; ++r) {
                    T v = acc[r];
                    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                    if (lane == 0 && rb + r < nrhs) ts[(j - j0) * nrhs + rb + r] = v;
                }
            }
        }
        load_block<T>(Lb, P, f, j0, j0, j1 - j0, j1 - j0);
        __syncthreads();
        for (int q = tid; q < (j1 - j0) * nrhs; q += 256) xs[q] = y[j0 * nrhs + q] - ts[q];
        __syncthreads();
        block_solve<T>(Lb, xs, j1 - j0, nrhs, unit, 0);
        __syncthreads();
        for (int q = tid; q < (j1 - j0) * nrhs; q += 256) y[j0 * nrhs + q] = xs[q];
        __syncthreads();
    }
    if (piv) {
        T* z = Z + (int64_t)c0 * nrhs;
        for (int q = tid; q < w * nrhs; q += 256) z[q] = y[q];
        __syncthreads();
        for (int q = tid; q < w * nrhs; q += 256) {
            const int kk = q / nrhs, r = q - kk * nrhs;
            Y[(int64_t)piv[c0 + kk] * nrhs + r] = z[q];
        }
    }
}

// ---------------------------------------------------------------------------
// large fronts: the sweeps of one front are split into NB-row diagonal blocks
// and run as level-batched launches with many workgroups per front.

// forward, before the block steps: zero the update vector, pull the children's update
// vectors (child_list order), then apply the local pivot order (piv) through Z.
// task: (s); one workgroup per front
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_pull(DevSym S, const Task* __restrict__ tasks, T* __restrict__ Y,
                                                 T* __restrict__ U, T* __restrict__ Z, int nrhs,
                                                 const int32_t* __restrict__ piv) {
    const int s = tasks[blockIdx.x].s;
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    T* us = S.uv_off[s] >= 0 ? U + S.uv_off[s] * nrhs : nullptr;
    if (us)
        for (int q = tid; q < m * nrhs; q += nt) us[q] = T(0);
    __syncthreads();
    for (int k = S.child_ptr[s]; k < S.child_ptr[s + 1]; ++k) {
        const int c = S.child_list[k];
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
    if (piv) {
        T* z = Z + (int64_t)c0 * nrhs;
        for (int q = tid; q < w * nrhs; q += nt) {
            const int kk = q / nrhs, r = q - kk * nrhs;
            z[q] = Y[(int64_t)piv[c0 + kk] * nrhs + r];
        }
        __syncthreads();
        for (int q = tid; q < w * nrhs; q += nt) Y[(int64_t)c0 * nrhs + q] = z[q];
    }
}

// forward block step: every workgroup solves the diagonal block [k, k + nb) (from Y) and
// applies it to its tile of 256 rows below: rows < w of Y, rows >= w of the update vector.
// The solved block goes to Z (written by the task with c = 1). task: (s, k, row0, c)
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_step(DevSym S, const Task* __restrict__ tasks,
                                                 const T* __restrict__ factor, T* __restrict__ Y, T* __restrict__ U,
                                                 T* __restrict__ Z, int nrhs, int unit) {
    __shared__ T Lb[NB][NB + 1];
    __shared__ T xs[NB * MAX_NRHS];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a, row0 = t.b;
    const int nb = min(NB, w - k);
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    load_block<T>(Lb, P, f, k, k, nb, nb);
    for (int q = tid; q < nb * nrhs; q += 256) xs[q] = Y[(int64_t)(c0 + k) * nrhs + q];
    __syncthreads();
    block_solve<T>(Lb, xs, nb, nrhs, unit, 1);
    __syncthreads();
    if (t.c)
        for (int q = tid; q < nb * nrhs; q += 256) Z[(int64_t)(c0 + k) * nrhs + q] = xs[q];
    const int i = row0 + tid;
    T* us = S.uv_off[s] >= 0 ? U + S.uv_off[s] * nrhs : nullptr;
    if (i < min(row0 + 256, f) && (i < w || us)) {
        for (int rb = 0; rb < nrhs; rb += 8) {
            const int nr = min(8, nrhs - rb);
            T acc[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] = T(0);
            for (int j0 = 0; j0 < nb; j0 += 16) {
                // sixteen panel entries of the row in flight together (clamped to the block)
                T l[16];
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) l[jj] = P[(int64_t)(k + min(j0 + jj, nb - 1)) * f + i];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) {
                    if (j0 + jj >= nb) break;
#pragma unroll
                    for (int r = 0; r < 8; ++r)
                        if (r < nr) acc[r] += l[jj] * xs[(j0 + jj) * nrhs + rb + r];
                }
            }
            rmw8<T>(i < w ? Y + (int64_t)(c0 + i) * nrhs + rb : us + (int64_t)(i - w) * nrhs + rb, acc, nr, i < w);
        }
    }
}

// after the block steps of a level: the solved values of the fronts' own columns from Z
// to Y (fwd), or to their pivot positions (bwd, piv). task: (s, col0)
template <typename T>
__global__ void __launch_bounds__(256) k_finish(DevSym S, const Task* __restrict__ tasks, T* __restrict__ Y,
                                               const T* __restrict__ Z, int nrhs, const int32_t* __restrict__ piv) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, w = S.fncols[s], c0 = S.super_ptr[s];
    const int j1 = min(t.a + 256, w);
    for (int q = threadIdx.x; q < (j1 - t.a) * nrhs; q += 256) {
        const int kk = t.a + q / nrhs, r = q % nrhs;
        const int dst = piv ? piv[c0 + kk] : c0 + kk;
        Y[(int64_t)dst * nrhs + r] = Z[(int64_t)(c0 + kk) * nrhs + r];
    }
}

// backward, before the block steps: y_j -= sum_{i >= w} L[i, j] y[rows[i]] for the columns
// [col0, col0 + 64) of the front, in two launches. k_bwd_l21: the partial sums over one chunk of
// L21_CHUNK rows below the columns. task: (s, col0, chunk, poff); the 64 x 64 tiles of L21 and
// the matching 64 rows of Y (gathered through rowval, 8 right-hand sides at a time) are staged
// in LDS, so L21 is read from memory once; thread (column jj, rhs pair) accumulates over the
// chunk's rows, into part[(poff + chunk w + col) nrhs + r]. k_bwd_l21_sum then subtracts the
// chunks' sums from Y in chunk order (a fixed order: the result does not depend on the schedule).
constexpr int L21_CHUNK = 256;

template <typename T>
__global__ void __launch_bounds__(256) k_bwd_l21(DevSym S, const Task* __restrict__ tasks,
                                                const T* __restrict__ factor, const T* __restrict__ Y,
                                                T* __restrict__ part, int nrhs) {
    __shared__ T Lt[64][65];  // Lt[jj][ii] = L(i0 + ii, col0 + jj)
    __shared__ T Yc[64][8];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], col0 = t.a;
    const int ncol = min(64, w - col0);
    const int r0 = w + t.b * L21_CHUNK, r1 = min(f, r0 + L21_CHUNK);
    const int tid = threadIdx.x, jj = tid & 63, rr = (tid >> 6) * 2;
    const T* P = factor + S.panel_off[s];
    const int32_t* rows = S.rowval + S.rowptr[s];
    T* pt = part + ((int64_t)t.c + (int64_t)t.b * w + col0) * nrhs;
    for (int rb = 0; rb < nrhs; rb += 8) {
        T acc0 = T(0), acc1 = T(0);
        for (int i0 = r0; i0 < r1; i0 += 64) {
            const int ni = min(64, r1 - i0);
            {
                // the tile's sixteen loads per thread and its Y rows in flight together
                // (clamped), then the masked LDS stores (zeros outside the tile)
                T v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int q = tid + 256 * u, c = q >> 6, ii = q & 63;
                    v[u] = P[(int64_t)(col0 + min(c, ncol - 1)) * f + i0 + min(ii, ni - 1)];
                }
                T yv[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int q = tid + 256 * u, ii = q >> 3, r = q & 7;
                    yv[u] = Y[(int64_t)rows[i0 + min(ii, ni - 1)] * nrhs + min(rb + r, nrhs - 1)];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int q = tid + 256 * u, c = q >> 6, ii = q & 63;
                    Lt[c][ii] = (c < ncol && ii < ni) ? v[u] : T(0);
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int q = tid + 256 * u, ii = q >> 3, r = q & 7;
                    Yc[ii][r] = (ii < ni && rb + r < nrhs) ? yv[u] : T(0);
                }
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
            if (rb + rr < nrhs) pt[(int64_t)jj * nrhs + rb + rr] = acc0;
            if (rb + rr + 1 < nrhs) pt[(int64_t)jj * nrhs + rb + rr + 1] = acc1;
        }
    }
}

// y_j -= sum over the front's chunks (in order) of the partial sums. task: (s, col0, nch, poff),
// the columns [col0, col0 + 64)
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_l21_sum(DevSym S, const Task* __restrict__ tasks, T* __restrict__ Y,
                                                    const T* __restrict__ part, int nrhs) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, w = S.fncols[s], col0 = t.a, nch = t.b;
    const int ncol = min(64, w - col0);
    T* y = Y + (int64_t)(S.super_ptr[s] + col0) * nrhs;
    const T* pt = part + ((int64_t)t.c + col0) * nrhs;
    for (int q = threadIdx.x; q < ncol * nrhs; q += 256) {
        T acc = pt[q];
        for (int c = 1; c < nch; ++c) acc += pt[(int64_t)c * w * nrhs + q];
        y[q] -= acc;
    }
}

// backward block step: every workgroup solves the diagonal block [k, k + nb) with L_bb^T
// (from Y) and applies it to its tile of 64 earlier columns [col0, col0 + 64) ∩ [0, k):
// y_j -= sum_{i in block} L[i, j] x_i. The solved block goes to Z (task with c = 1).
// task: (s, k, col0, c)
template <typename T>
__global__ void __launch_bounds__(256) k_bwd_step(DevSym S, const Task* __restrict__ tasks,
                                                 const T* __restrict__ factor, T* __restrict__ Y, T* __restrict__ Z,
                                                 int nrhs, int unit) {
    __shared__ T Lb[NB][NB + 1];
    __shared__ T Lt[NB][NB + 1];  // Lt[jj][i] = L(k + i, col0 + jj)
    __shared__ T xs[NB * MAX_NRHS];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a, col0 = t.b;
    const int nb = min(NB, w - k);
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    const int ncol = max(0, min(col0 + 64, k) - col0);
    load_block<T>(Lb, P, f, k, k, nb, nb);
    load_block<T>(Lt, P, f, k, col0, nb, ncol);
    for (int q = tid; q < nb * nrhs; q += 256) xs[q] = Y[(int64_t)(c0 + k) * nrhs + q];
    __syncthreads();
    block_solve<T>(Lb, xs, nb, nrhs, unit, 0);
    __syncthreads();
    if (t.c)
        for (int q = tid; q < nb * nrhs; q += 256) Z[(int64_t)(c0 + k) * nrhs + q] = xs[q];
    const int jj = tid & 63;
    if (jj < ncol)
        for (int r = tid >> 6; r < nrhs; r += 4) {
            T acc = T(0);
#pragma unroll 16
            for (int i = 0; i < nb; ++i) acc += Lt[jj][i] * xs[i * nrhs + r];
            Y[(int64_t)(c0 + col0 + jj) * nrhs + r] -= acc;
        }
}

// Y <- D^-1 Y with 1x1 and 2x2 blocks (src/solve/sweeps.jl:549-570); one thread per (row, rhs)
template <typename T>
__global__ void k_diag(int n, int nrhs, const T* __restrict__ d, const int8_t* __restrict__ kind, T* __restrict__ Y) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / nrhs), r = (int)(q - (int64_t)k * nrhs);
        const int8_t kd = kind ? kind[k] : KIND_1X1;
        if (kd == KIND_2X2_SECOND) continue;
        if (kd == KIND_2X2_FIRST) {
            const T a = d[k], b = d[n + k], c = d[k + 1];
            const T det = a * c - b * b;
            const T y1 = Y[(int64_t)k * nrhs + r], y2 = Y[(int64_t)(k + 1) * nrhs + r];
            Y[(int64_t)k * nrhs + r] = (c * y1 - b * y2) / det;
            Y[(int64_t)(k + 1) * nrhs + r] = (a * y2 - b * y1) / det;
        } else {
            Y[q] /= d[k];
        }
    }
}

// R = B - M X over the full-pattern map (rowptr/colval/src 0-based; src < 0: -(p + 1), conjugate),
// FP64; one thread per (row, rhs); `roff` selects the rows of M^T of a "G" map. B = nullptr gives
// the product R = M X (the operator of FGMRES refinement).
__global__ void k_residual(int n, int nrhs, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ colval,
                           const int64_t* __restrict__ src, const double* __restrict__ nz, const double* __restrict__ X,
                           int64_t ldx, const double* __restrict__ B, int64_t ldb, double* __restrict__ R) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(q / n), i = (int)(q - (int64_t)r * n);
        double acc = 0.0;
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            const int64_t p = src[e];
            const double v = nz[p >= 0 ? p : -(p + 1)];
            acc += v * X[colval[e] + r * ldx];
        }
        R[q] = B ? B[i + r * ldb] - acc : acc;
    }
}

// ---------------------------------------------------------------------------
// FGMRES refinement: vector kernels over the n nrhs stacked columns, FP64. Reductions run in a
// fixed order (DOT_BLOCKS grid-stride partial sums, then one workgroup's tree), so every run
// gives the same bits.

__device__ inline double block_sum256(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    return red[0];
}

// part[block] = partial sum of x_i y_i
__global__ void __launch_bounds__(256) k_dot_partial(int64_t len, const double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < len; i += (int64_t)gridDim.x * 256)
        acc = fma(x[i], y[i], acc);
    const double v = block_sum256(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// out[0] = sum of part[0, np), np <= 256
__global__ void __launch_bounds__(256) k_dot_final(int np, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double red[256];
    const double v = block_sum256(threadIdx.x < np ? part[threadIdx.x] : 0.0, red);
    if (threadIdx.x == 0) out[0] = v;
}

// y -= a[0] x (the Gram-Schmidt coefficient stays on the device)
__global__ void k_axpy_neg(int64_t len, double* __restrict__ y, const double* __restrict__ x,
                           const double* __restrict__ a) {
    const double s = -a[0];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = fma(s, x[i], y[i]);
}

__global__ void k_scale(int64_t len, double* __restrict__ y, double s) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        y[i] *= s;
}

// X(:, r) += sum_j c_j Z_j(:, r), the Z_j stacked n nrhs vectors at stride zs
__global__ void k_add_combination(int n, int nrhs, double* __restrict__ X, int64_t ldx, const double* __restrict__ Z,
                                  int64_t zs, int k, const double* __restrict__ c) {
    const int64_t total = (int64_t)n * nrhs;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < total; q += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(q / n), i = (int)(q - (int64_t)r * n);
        double acc = 0.0;
        for (int j = 0; j < k; ++j) acc = fma(c[j], Z[j * zs + q], acc);
        X[i + r * ldx] += acc;
    }
}

// Adjoint of x = A^-1 b with respect to the stored values: grad[p] = -sum_r lambda_i x_j over
// the right-hand sides r, for the entry p at (i, j); a mirrored entry of a symmetric structure
// stands for (i, j) and (j, i) and gets both terms; an entry the solver ignores gets 0.
// One thread per stored entry, no atomics.
__global__ void k_adjoint_scatter(int64_t nnz, int nrhs, const int32_t* __restrict__ ai, const int32_t* __restrict__ aj,
                                  const int8_t* __restrict__ aflag, const double* __restrict__ lam, int64_t ldl,
                                  const double* __restrict__ X, int64_t ldx, double* __restrict__ grad) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const int8_t fl = aflag[p];
        double g = 0.0;
        if (fl) {
            const int i = ai[p], j = aj[p];
            for (int r = 0; r < nrhs; ++r) {
                g += lam[i + r * ldl] * X[j + r * ldx];
                if (fl == 2) g += lam[j + r * ldl] * X[i + r * ldx];
            }
        }
        grad[p] = -g;
    }
}

// ---------------------------------------------------------------------------

static int grid_for(int64_t total) { return (int)std::min<int64_t>(std::max<int64_t>((total + 255) / 256, 1), 4096); }

template <typename T>
void launch_permute_in(hipStream_t st, int n, int nrhs, const int32_t* perm, const double* B, int64_t ldb, T* Y,
                       const int32_t* cq, const double* sc) {
    hipLaunchKernelGGL(k_permute_in<T>, dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, n, nrhs, perm, cq, sc, B,
                       ldb, Y);
}

void launch_scale_values(hipStream_t st, int64_t nnz, const double* w, const double* nz, double* out) {
    hipLaunchKernelGGL(k_scale_values, dim3(grid_for(nnz)), dim3(256), 0, st, nnz, w, nz, out);
}

template <typename T>
void launch_schur_pull(hipStream_t st, const DevSym& S, int s, T* Y, const T* U, int nrhs) {
    hipLaunchKernelGGL(k_schur_pull<T>, dim3(1), dim3(256), 0, st, S, s, Y, U, nrhs);
}

template <typename T>
void launch_permute_out(hipStream_t st, int n, int nrhs, const int32_t* perm, const T* Y, double* X, int64_t ldx,
                        int add, const int32_t* cq, const double* sc) {
    hipLaunchKernelGGL(k_permute_out<T>, dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, n, nrhs, perm, cq, sc, Y,
                       X, ldx, add);
}

template <typename T>
void launch_fwd(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, const T* factor, T* Y, T* U, T* Z,
                int nrhs, int unit, const int32_t* piv) {
    hipLaunchKernelGGL(k_fwd<T>, dim3(nfronts), dim3(256), 0, st, S, fronts, factor, Y, U, Z, nrhs, unit, piv);
}

template <typename T>
void launch_bwd(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, const T* factor, T* Y, T* Z,
                int nrhs, int unit, const int32_t* piv) {
    size_t lds = (size_t)128 * nrhs * sizeof(T);
    hipLaunchKernelGGL((k_bwd<T, 8>), dim3(nfronts), dim3(256), lds, st, S, fronts, factor, Y, Z, nrhs, unit, piv);
}

template <typename T>
void launch_fwd_pull(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* Y, T* U, T* Z, int nrhs,
                     const int32_t* piv) {
    hipLaunchKernelGGL(k_fwd_pull<T>, dim3(n), dim3(256), 0, st, S, tasks, Y, U, Z, nrhs, piv);
}

template <typename T>
void launch_fwd_step(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, T* Y, T* U, T* Z,
                     int nrhs, int unit) {
    hipLaunchKernelGGL(k_fwd_step<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, Y, U, Z, nrhs, unit);
}

template <typename T>
void launch_finish(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* Y, const T* Z, int nrhs,
                   const int32_t* piv) {
    hipLaunchKernelGGL(k_finish<T>, dim3(n), dim3(256), 0, st, S, tasks, Y, Z, nrhs, piv);
}

template <typename T>
void launch_bwd_l21(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* Y, T* part,
                    int nrhs) {
    hipLaunchKernelGGL(k_bwd_l21<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, Y, part, nrhs);
}

template <typename T>
void launch_bwd_l21_sum(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* Y, const T* part, int nrhs) {
    hipLaunchKernelGGL(k_bwd_l21_sum<T>, dim3(n), dim3(256), 0, st, S, tasks, Y, part, nrhs);
}

int bwd_l21_chunk() { return L21_CHUNK; }

template <typename T>
void launch_bwd_step(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, T* Y, T* Z, int nrhs,
                     int unit) {
    hipLaunchKernelGGL(k_bwd_step<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, Y, Z, nrhs, unit);
}

template <typename T>
void launch_diag(hipStream_t st, int n, int nrhs, const T* d, const int8_t* kind, T* Y) {
    hipLaunchKernelGGL(k_diag<T>, dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, n, nrhs, d, kind, Y);
}

void launch_adjoint_scatter(hipStream_t st, int64_t nnz, int nrhs, const int32_t* ai, const int32_t* aj,
                            const int8_t* aflag, const double* lam, int64_t ldl, const double* X, int64_t ldx,
                            double* grad) {
    hipLaunchKernelGGL(k_adjoint_scatter, dim3(grid_for(nnz)), dim3(256), 0, st, nnz, nrhs, ai, aj, aflag, lam, ldl, X,
                       ldx, grad);
}

void launch_dot(hipStream_t st, int64_t len, const double* x, const double* y, double* part, double* out) {
    hipLaunchKernelGGL(k_dot_partial, dim3(DOT_BLOCKS), dim3(256), 0, st, len, x, y, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, st, DOT_BLOCKS, part, out);
}

void launch_axpy_neg(hipStream_t st, int64_t len, double* y, const double* x, const double* a) {
    hipLaunchKernelGGL(k_axpy_neg, dim3(grid_for(len)), dim3(256), 0, st, len, y, x, a);
}

void launch_scale(hipStream_t st, int64_t len, double* y, double s) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(len)), dim3(256), 0, st, len, y, s);
}

void launch_add_combination(hipStream_t st, int n, int nrhs, double* X, int64_t ldx, const double* Z, int64_t zs, int k,
                            const double* c) {
    hipLaunchKernelGGL(k_add_combination, dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, n, nrhs, X, ldx, Z, zs,
                       k, c);
}

void launch_residual(hipStream_t st, int n, int nrhs, const int64_t* rowptr, const int32_t* colval, const int64_t* src,
                     const double* nz, const double* X, int64_t ldx, const double* B, int64_t ldb, double* R) {
    hipLaunchKernelGGL(k_residual, dim3(grid_for((int64_t)n * nrhs)), dim3(256), 0, st, n, nrhs, rowptr, colval, src,
                       nz, X, ldx, B, ldb, R);
}

#define INST(T)                                                                                                     \
    template void launch_permute_in<T>(hipStream_t, int, int, const int32_t*, const double*, int64_t, T*,          \
                                       const int32_t*, const double*);                                             \
    template void launch_schur_pull<T>(hipStream_t, const DevSym&, int, T*, const T*, int);                        \
    template void launch_permute_out<T>(hipStream_t, int, int, const int32_t*, const T*, double*, int64_t, int,     \
                                        const int32_t*, const double*);                                             \
    template void launch_fwd<T>(hipStream_t, const DevSym&, const int32_t*, int, const T*, T*, T*, T*, int, int,    \
                                const int32_t*);                                                                    \
    template void launch_bwd<T>(hipStream_t, const DevSym&, const int32_t*, int, const T*, T*, T*, int, int,        \
                                const int32_t*);                                                                    \
    template void launch_diag<T>(hipStream_t, int, int, const T*, const int8_t*, T*);                              \
    template void launch_fwd_pull<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, T*, int, const int32_t*); \
    template void launch_fwd_step<T>(hipStream_t, const DevSym&, const Task*, int, const T*, T*, T*, T*, int, int);  \
    template void launch_finish<T>(hipStream_t, const DevSym&, const Task*, int, T*, const T*, int, const int32_t*); \
    template void launch_bwd_l21<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, int);     \
    template void launch_bwd_l21_sum<T>(hipStream_t, const DevSym&, const Task*, int, T*, const T*, int);          \
    template void launch_bwd_step<T>(hipStream_t, const DevSym&, const Task*, int, const T*, T*, T*, int, int);
INST(double)
INST(float)
#undef INST

}  // namespace sds
