// Cholesky sweeps with inverted diagonal blocks (gfx950).
//
// After a Cholesky factorization one launch inverts every NB x NB diagonal block of every
// front (Linv = L_bb^-1, lower triangular, kept in the DB buffer: NB entries per factor column,
// column-major with leading dimension NB), and a second one forms the inverses of the wider
// solve steps' diagonal blocks from them. The sweeps then need no serial substitution: every
// triangular solve is a GEMV with an inverse, and every launch of a level is fully parallel
// across rows or columns. This is the partitioned-inverse form of the reference's
// solve_alg = "algo1" (src/solver.jl:834, src/solve/fused.jl). Every entry is produced by
// one thread or one fixed reduction tree: no atomics, and the result is bitwise reproducible.
//
// A front's fully summed columns are cut into solve steps of at most SOLVE_STEP columns
// [c0, c1) (one step unless the front is wider). A narrow step (at most NB columns) uses the
// inverse Linv of its single diagonal block; a wide step uses X, the inverse of its whole
// diagonal block, formed after the factorization by k_step_inverse. Per tree level and step:
//   forward : pull the children's update vectors (at a front's first step; each row gathers its
//             entries, children in child_list order, from the host's pull map) and form
//             y_step = inverse * t_step:
//             k_fwd_diag (narrow, one workgroup per step; a narrow front with at most FUSE_ROWS
//             rows below its columns forms its update vector there too) or k_fwd_inv (wide,
//             16 rows per workgroup, into Z); then k_fwd_rows adds L(rows, step) y_step to the
//             rows below the step, 64 per workgroup (the update vector rows start from the
//             children's entries of their rows; a wide step's y_step is read from Z and copied
//             to y by one of its tasks, or by k_copy_rows when it has no rows below).
//   backward: k_bwd_l21t, partial sums of L(rows below the step, step)^T x per chunk of
//             L21T_CHUNK rows (one wave per 8 columns, lanes along the rows, a reduce-scatter
//             across the lanes at the end); then x_step = inverse^T (y_step - partials):
//             k_bwd_diag (narrow) or k_sub_partials (y_step less the partials, into Z) +
//             k_bwd_inv (wide, one wave per 8 columns, from Z into y).
#include "sds_common.h"
#include "sds_kernels.h"

namespace sds {

// the solve kernels contract multiply-adds (their results are compared with a tolerance)
#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// Linv of the diagonal block [k, k + nb) of a small front: BS lanes per block, 64 / BS blocks
// per wave (nb <= BS); lane c computes column c of the inverse by forward substitution with the
// partial sums in registers (fully unrolled: the block is padded with the identity past nb, so
// no guards touch the register array; no scheduling barriers either: with them the compiler
// spills this kernel). The padding leaves the entries of the block itself unchanged, so every
// BS gives the same Linv bit for bit; the host sends each block to the smallest BS that holds
// it (most small fronts are under 16 columns wide). The large fronts' blocks are inverted by
// k_big_diag during the factorization. Linv goes to DB (NB rows per column, zero past nb).
// task: (s, k); a task with c = 1 takes L from DB and writes it back to the panel.

template <typename T, int BS>
__global__ void __launch_bounds__(64) k_block_inverse(DevSym S, const Task* __restrict__ tasks, int n,
                                                     T* __restrict__ factor, T* __restrict__ DB) {
    constexpr int G = 64 / BS;
    __shared__ T Lb[G][BS][BS + 1];  // Lb[g][m][r] = L(k + r, k + m) of group g's block
    __shared__ T rinv[G][BS];
    const int g = threadIdx.x / BS, c = threadIdx.x % BS;
    const int id = blockIdx.x * G + g;
    const bool live = id < n;
    const Task t = tasks[live ? id : n - 1];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a;
    const int nb = min(BS, w - k);
    T* P = factor + S.panel_off[s];
    T* blk = DB + (int64_t)(S.super_ptr[s] + k) * NB;
    {
        // the column's BS loads issued together (clamped into the block), then the masking
        const T* src = t.c ? blk + min(c, nb - 1) : P + (int64_t)k * f + k + min(c, nb - 1);
        const int64_t ld = t.c ? NB : f;
        T lv[BS];
#pragma unroll
        for (int m = 0; m < BS; ++m) lv[m] = src[(int64_t)min(m, nb - 1) * ld];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int m = 0; m < BS; ++m) {
            const int r = c;
            Lb[g][m][r] = m < nb && r < nb ? (r >= m ? lv[m] : T(0)) : (r == m ? T(1) : T(0));
        }
    }
    __syncthreads();
    if (live && t.c && c < nb)
        for (int m = 0; m <= c; ++m) P[(int64_t)(k + m) * f + k + c] = Lb[g][m][c];
    rinv[g][c] = T(1) / Lb[g][c][c];
    __syncthreads();
    T x[BS];
#pragma unroll
    for (int r = 0; r < BS; ++r) x[r] = r == c ? T(1) : T(0);
#pragma unroll
    for (int m = 0; m < BS; ++m) {
        x[m] = x[m] * rinv[g][m];
        const T xm = x[m];
#pragma unroll
        for (int r = m + 1; r < BS; ++r) x[r] = fma(-Lb[g][m][r], xm, x[r]);
    }
    // column c of the block (only the front's own columns: the last block may be narrow)
    if (live && c < nb) {
        T* out = blk + (int64_t)c * NB;
#pragma unroll
        for (int r = 0; r < BS; ++r) out[r] = r < nb ? x[r] : T(0);
#pragma unroll
        for (int r = BS; r < NB; ++r) out[r] = T(0);
    }
}

// ---------------------------------------------------------------------------
// helpers

extern __shared__ __align__(16) unsigned char sweep_lds[];



// one recursive-halving exchange at lane distance D: the lanes with bit D set keep the upper
// halves a[D, 2D), the others the lower, each adding its partner's copy
template <int D, typename T>
__device__ __attribute__((always_inline)) inline void halving_step(T (&a)[64], int lane) {
    const bool up = (lane & D) != 0;
#pragma unroll
    for (int v = 0; v < D; ++v) {
        const T send = up ? a[v] : a[v + D];
        const T keep = up ? a[v + D] : a[v];
        a[v] = keep + __shfl_xor(send, D);
    }
}

// sum over the 64 lanes of a[v] for v < 64, reduce-scattered: lane l returns the sum of a[l]
// (recursive halving, 63 exchanges; a fixed summation order)
template <typename T>
__device__ __attribute__((always_inline)) inline T wave_reduce_scatter64(T (&a)[64]) {
    const int lane = threadIdx.x & 63;
    halving_step<32>(a, lane);
    halving_step<16>(a, lane);
    halving_step<8>(a, lane);
    halving_step<4>(a, lane);
    halving_step<2>(a, lane);
    halving_step<1>(a, lane);
    return a[0];
}

// ---------------------------------------------------------------------------
// forward

// one workgroup per narrow step (at most NB columns: a whole narrow front, or the last columns
// [c0, c1) of a wide front): task (step, fused). At c0 = 0 pull the children's update vectors;
// y = Linv t over the step's columns (entry (j, r) per thread, Linv rows read
// coalesced, t in LDS); a fused front (at most FUSE_ROWS rows below) also forms its update
// vector u_i = L(i, front) y, one row per thread. LDS: 2 NB 8 entries.
template <typename T>
__device__ __attribute__((always_inline)) inline void fwd_diag_body(const DevSym& S,
                                                                    const SolveStep* __restrict__ steps, const Task tk,
                                                                    const T* __restrict__ factor,
                                                                    const T* __restrict__ DB, T* __restrict__ Y,
                                                                    T* __restrict__ U, int nrhs) {
    T* tb = reinterpret_cast<T*>(sweep_lds);  // NB x 8: t
    T* yx = tb + NB * 8;                       // NB x 8: y
    const int tid = threadIdx.x, nt = blockDim.x;
    const SolveStep st = steps[tk.s];
    const int s = st.s, c0 = st.c0, nw = st.c1 - c0;
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int cg = S.super_ptr[s];
    const T* P = factor + S.panel_off[s];
    const bool fused = tk.a != 0;  // only a whole narrow front (c0 = 0, nw = w) is fused
    T* us = fused && m > 0 && S.uv_off[s] >= 0 ? U + S.uv_off[s] * nrhs : nullptr;
    // the children's update vectors at the front's first step, row by row from the pull map
    // (child_list order; only the rows < w unless fused: the update vector rows start from zero).
    // The last narrow step of a wide front (c0 > 0) got its rows' entries from k_fwd_rows at the
    // first step
    if (c0 == 0) {
        const int32_t* pp = S.pm_ptr + S.pm_row[s];
        const int rows = us ? f : w;
        for (int q = tid; q < rows * nrhs; q += nt) {
            const int i = q / nrhs, r = q - i * nrhs;
            const int e0 = pp[i], e1 = pp[i + 1];
            if (i < w) {
                if (e0 == e1) continue;
                T* yp = Y + (int64_t)(cg + i) * nrhs + r;
                T v = *yp;
                for (int e = e0; e < e1; ++e) v -= U[(int64_t)S.pm_src[e] * nrhs + r];
                *yp = v;
            } else {
                T v = T(0);
                for (int e = e0; e < e1; ++e) v += U[(int64_t)S.pm_src[e] * nrhs + r];
                us[(i - w) * nrhs + r] = v;
            }
        }
    }
    __syncthreads();
    T* y = Y + (int64_t)(cg + c0) * nrhs;
    const T* Li = DB + (int64_t)(cg + c0) * NB;
    for (int rb = 0; rb < nrhs; rb += 8) {
        const int nr = min(8, nrhs - rb);
        for (int q = tid; q < nw * 8; q += nt) tb[q] = (q & 7) < nr ? y[(q >> 3) * nrhs + rb + (q & 7)] : T(0);
        __syncthreads();
        for (int q = tid; q < nw * 8; q += nt) {
            const int j = q >> 3, r = q & 7;
            T acc = T(0);
#pragma unroll 8
            for (int i = 0; i <= j; ++i) acc += Li[(int64_t)i * NB + j] * tb[i * 8 + r];
            yx[q] = acc;
            if (r < nr) y[j * nrhs + rb + r] = acc;
        }
        __syncthreads();
        if (us)
            for (int i = w + tid; i < f; i += nt) {
                T a[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) a[r] = T(0);
#pragma unroll 8
                for (int j = 0; j < nw; ++j) {
                    const T lv = P[(int64_t)j * f + i];
#pragma unroll
                    for (int r = 0; r < 8; ++r) a[r] += lv * yx[j * 8 + r];
                }
                rmw8<T>(us + (i - w) * nrhs + rb, a, nr, false);
            }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fwd_diag(DevSym S, const SolveStep* __restrict__ steps,
                                                 const Task* __restrict__ tasks, const T* __restrict__ factor,
                                                 const T* __restrict__ DB, T* __restrict__ Y, T* __restrict__ U,
                                                 int nrhs) {
    fwd_diag_body<T>(S, steps, tasks[blockIdx.x], factor, DB, Y, U, nrhs);
}

// rows [r0, r1) below a front step (64 per workgroup of 8 waves, within L11 or within the update
// vector): y_i -= L(i, step) y_step, or u_i += L(i, step) y_step. At the front's first step the
// rows get the children's entries of those rows first (child_list order; update vector rows are
// zeroed before). Wave v takes the step's 16-column batches v, v + 8, ...; y_step is staged in
// LDS; the eight partial sums are added in wave order. task: (step, r0, copy, wide): a wide
// step's y_step is read from Z, where k_fwd_inv left it, and the task with copy = 1 also copies it
// to Y (no task of the launch reads those rows of Y).
// LDS: SOLVE_STEP 8 + 8 64 8 entries.
template <typename T>
__global__ void __launch_bounds__(512) k_fwd_rows(DevSym S, const SolveStep* __restrict__ steps,
                                                 const Task* __restrict__ tasks, const T* __restrict__ factor,
                                                 T* __restrict__ Y, const T* __restrict__ Z, T* __restrict__ U,
                                                 int nrhs) {
    constexpr int NWV = 8;
    T* ys = reinterpret_cast<T*>(sweep_lds);  // step columns x 8
    T* red = ys + SOLVE_STEP * 8;              // NWV waves x 64 rows x 8
    const Task tk = tasks[blockIdx.x];
    const SolveStep st = steps[tk.s];
    const int s = st.s, c0 = st.c0, c1 = st.c1;
    const int f = S.fnrows[s], w = S.fncols[s];
    const int cg = S.super_ptr[s];
    const int r0 = tk.a, r1 = min(r0 + 64, r0 < w ? w : f);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nt = NWV * 64;
    const T* P = factor + S.panel_off[s];
    const bool below = r0 >= w;
    T* us = below ? U + S.uv_off[s] * nrhs : nullptr;
    if (c0 == 0) {
        // the rows' entries of the children's update vectors, row by row from the pull map
        // (child_list order; update vector rows start from zero)
        const int32_t* pp = S.pm_ptr + S.pm_row[s];
        for (int q = tid; q < (r1 - r0) * nrhs; q += nt) {
            const int di = q / nrhs, i = r0 + di, r = q - di * nrhs;
            const int e0 = pp[i], e1 = pp[i + 1];
            if (below) {
                T v = T(0);
                for (int e = e0; e < e1; ++e) v += U[(int64_t)S.pm_src[e] * nrhs + r];
                us[(int64_t)(i - w) * nrhs + r] = v;
            } else if (e0 < e1) {
                T* yp = Y + (int64_t)(cg + i) * nrhs + r;
                T v = *yp;
                for (int e = e0; e < e1; ++e) v -= U[(int64_t)S.pm_src[e] * nrhs + r];
                *yp = v;
            }
        }
        __syncthreads();
    }
    const int nw = c1 - c0;
    const int i = min(r0 + lane, r1 - 1);
    const T* pi = P + i;
    const T* ysrc = tk.c ? Z : Y;
    if (tk.b)
        for (int q = tid; q < nw * nrhs; q += nt) Y[(int64_t)(cg + c0) * nrhs + q] = Z[(int64_t)(cg + c0) * nrhs + q];
    for (int rb = 0; rb < nrhs; rb += 8) {
        const int nr = min(8, nrhs - rb);
        for (int q = tid; q < nw * 8; q += nt) {
            const int j = q >> 3, r = q & 7;
            ys[q] = r < nr ? ysrc[(int64_t)(cg + c0 + j) * nrhs + rb + r] : T(0);
        }
        __syncthreads();
        T acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = T(0);
        for (int jb = 16 * wv; jb < nw; jb += 16 * NWV) {
            T l[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj)  // the factor streams through once per sweep
                l[jj] = __builtin_nontemporal_load(pi + (int64_t)(c0 + min(jb + jj, nw - 1)) * f);
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const T lv = jb + jj < nw ? l[jj] : T(0);
                const T* yj = ys + min(jb + jj, nw - 1) * 8;
#pragma unroll
                for (int r = 0; r < 8; ++r) acc[r] += lv * yj[r];
            }
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) red[(wv * 64 + lane) * 8 + r] = acc[r];
        __syncthreads();
        if (wv == 0 && r0 + lane < r1) {
            T v[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                v[r] = red[lane * 8 + r];
#pragma unroll
                for (int u = 1; u < NWV; ++u) v[r] += red[(u * 64 + lane) * 8 + r];
            }
            rmw8<T>(below ? us + (int64_t)(i - w) * nrhs + rb : Y + (int64_t)(cg + i) * nrhs + rb, v, nr, !below);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// backward

// partial sums of L(rows below the step, step)^T x over the chunk of L21T_CHUNK rows of a task
// (step, col0, chunk): one wave for the 8 columns from col0; lanes along the rows; 8
// right-hand sides per pass. The 64 sums are reduce-scattered across the lanes (lane l: column
// l / 8, right-hand side l % 8). part[(poff + chunk nw + j - c0) nrhs + r], nw = c1 - c0.
template <typename T>
__global__ void __launch_bounds__(64) k_bwd_l21t(DevSym S, const SolveStep* __restrict__ steps,
                                                const Task* __restrict__ tasks, const T* __restrict__ factor,
                                                const T* __restrict__ Y, T* __restrict__ part, int nrhs) {
    const Task tk = tasks[blockIdx.x];
    const SolveStep st = steps[tk.s];
    const int s = st.s, c0 = st.c0, c1 = st.c1, chunk = tk.b;
    const int col0 = tk.a;
    const int f = S.fnrows[s];
    const int lane = threadIdx.x;
    const T* P = factor + S.panel_off[s];
    const int32_t* rows = S.rowval + S.rowptr[s];
    const int r0 = c1 + chunk * L21T_CHUNK, r1 = min(f, r0 + L21T_CHUNK);
    T* out = part + (st.poff + (int64_t)chunk * (c1 - c0)) * nrhs;
    for (int rb = 0; rb < nrhs; rb += 8) {
        T a[64];
#pragma unroll
        for (int v = 0; v < 64; ++v) a[v] = T(0);
#pragma unroll 2
        for (int i = r0 + lane; i < r1; i += 64) {
            const T* yr = Y + (int64_t)rows[i] * nrhs;
            T x[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) x[r] = yr[min(rb + r, nrhs - 1)];
            const T* pc = P + (int64_t)col0 * f + i;  // column col0 + c; columns past c1 read c1 - 1
            T l[8];  // the row's eight loads issued together (the FP32 build otherwise waits per load)
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                l[c] = __builtin_nontemporal_load(pc);  // the factor streams through once per sweep
                pc += col0 + c + 1 < c1 ? f : 0;
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int r = 0; r < 8; ++r) a[c * 8 + r] += l[c] * x[r];
        }
        const T v = wave_reduce_scatter64<T>(a);
        const int c = lane >> 3, r = lane & 7;
        if (col0 + c < c1 && rb + r < nrhs) out[(int64_t)(col0 + c - c0) * nrhs + rb + r] = v;
    }
}

// one workgroup per narrow step (a whole narrow front, or the last columns [c0, c1) of a wide
// front): task (step, 0). t = y - the L21^T x partial sums (chunk order), x = Linv^T t over the
// step's columns (entry (j, r) per thread). LDS: NB 8 entries.
template <typename T>
__device__ __attribute__((always_inline)) inline void bwd_diag_body(const DevSym& S,
                                                                    const SolveStep* __restrict__ steps, const Task tk,
                                                                    const T* __restrict__ factor,
                                                                    const T* __restrict__ DB, T* __restrict__ Y,
                                                                    const T* __restrict__ part, int nrhs) {
    T* tb = reinterpret_cast<T*>(sweep_lds);  // NB x 8
    const int tid = threadIdx.x, nt = blockDim.x;
    const SolveStep st = steps[tk.s];
    const int s = st.s, c0 = st.c0, nw = st.c1 - c0;
    const int cg = S.super_ptr[s];
    T* y = Y + (int64_t)(cg + c0) * nrhs;
    const T* pp = part + st.poff * nrhs;
    const T* Li = DB + (int64_t)(cg + c0) * NB;
    for (int rb = 0; rb < nrhs; rb += 8) {
        const int nr = min(8, nrhs - rb);
        for (int q = tid; q < nw * 8; q += nt) {
            const int j = q >> 3, r = q & 7;
            T v = T(0);
            if (r < nr) {
                v = y[j * nrhs + rb + r];
                for (int ch = 0; ch < st.nch; ++ch) v -= pp[((int64_t)ch * nw + j) * nrhs + rb + r];
            }
            tb[q] = v;
        }
        __syncthreads();
        for (int q = tid; q < nw * 8; q += nt) {
            const int j = q >> 3, r = q & 7;
            T acc = T(0);
#pragma unroll 8
            for (int i = j; i < nw; ++i) acc += Li[(int64_t)j * NB + i] * tb[i * 8 + r];
            if (r < nr) y[j * nrhs + rb + r] = acc;
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bwd_diag(DevSym S, const SolveStep* __restrict__ steps,
                                                 const Task* __restrict__ tasks, const T* __restrict__ factor,
                                                 const T* __restrict__ DB, T* __restrict__ Y,
                                                 const T* __restrict__ part, int nrhs) {
    bwd_diag_body<T>(S, steps, tasks[blockIdx.x], factor, DB, Y, part, nrhs);
}

// ---------------------------------------------------------------------------
// wide solve steps (more than NB columns): the inverse X of the step's diagonal block

// X = L(step, step)^-1 for a wide step, one workgroup per (step, column block kk): X_kk = Linv_kk
// (from DB), then for the blocks below X_ik = -Linv_ii sum_{j = kk}^{i - 1} L_ij X_jk, 64 x 64 x 64
// products through LDS on the matrix cores (wave w: the 32 x 32 quadrant (w / 2, w % 2) of the
// 64 x 64 result, four 16 x 16 accumulators). X column-major, ld nw, at xoff. The operand tiles
// are loaded into registers in one unrolled batch (16 loads in flight per thread), and the next
// product's tiles are loaded while the current one is computed: the chain of kk = 0 in a 512-wide
// step is 35 products long, so each product's load latency is on the critical path.
// task: (step, kk)
template <typename T>
__device__ __attribute__((always_inline)) inline void tile_product_mfma(const T (*A)[NB + 1], const T (*B)[NB + 1],
                                                                        typename Mfma<T>::acc_t (&acc)[2][2],
                                                                        int row0, int col0, int l16, int l4) {
#pragma unroll 4
    for (int kk = 0; kk < NB / 4; ++kk) {
        T av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) av[a] = A[row0 + 16 * a + l16][4 * kk + l4];
#pragma unroll
        for (int b = 0; b < 2; ++b) bv[b] = B[4 * kk + l4][col0 + 16 * b + l16];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = Mfma<T>::mac(av[a], bv[b], acc[a][b]);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_step_inverse(DevSym S, const SolveStep* __restrict__ steps,
                                                     const Task* __restrict__ tasks, const T* __restrict__ factor,
                                                     const T* __restrict__ DB, T* __restrict__ XB) {
    typedef typename Mfma<T>::acc_t acc_t;
    constexpr int PER = NB * NB / 256;  // tile entries per thread: row tid % NB, columns tid / NB + 4 i
    __shared__ T A[NB][NB + 1];
    __shared__ T B[NB][NB + 1];
    const Task tk = tasks[blockIdx.x];
    const SolveStep st = steps[tk.s];
    const int s = st.s, c0 = st.c0, c1 = st.c1, nw = c1 - c0, nblk = (nw + NB - 1) / NB;
    const int f = S.fnrows[s], cg = S.super_ptr[s];
    const T* P = factor + S.panel_off[s];
    T* X = XB + st.xoff;
    const int kc = tk.a * NB, nbk = min(NB, nw - kc);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int row0 = 32 * (wv >> 1), col0 = 32 * (wv & 1);
    const int lr = tid % NB, lc = tid / NB;
    {
        // X_kk = Linv_kk: the loads issued together (clamped into the block), then the stores
        T xv[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i)
            xv[i] = DB[(int64_t)(cg + c0 + kc + min(lc + 4 * i, nbk - 1)) * NB + min(lr, nbk - 1)];
#pragma unroll
        for (int i = 0; i < PER; ++i) asm volatile("" : "+v"(xv[i]));
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int c = lc + 4 * i;
            if (c < nbk && lr < nbk) X[(int64_t)(kc + c) * nw + kc + lr] = xv[i];
        }
    }
    // the operands of product (ib, jb): A = L(ic + r, jc + c), B = X(jc + r, kc + c); loaded from
    // clamped in-bounds addresses (jb < ib < nblk: block jb is full), masked at the use
    T ra[PER], rb[PER];
    int nbl = NB;  // rows of block ib of the loaded operands
    auto load = [&](int ib, int jb) {
        const int ic = ib * NB, jc = jb * NB;
        nbl = min(NB, nw - ic);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int c = lc + 4 * i;
            ra[i] = P[(int64_t)(c0 + jc + c) * f + c0 + ic + min(lr, nbl - 1)];
            rb[i] = X[(int64_t)(kc + min(c, nbk - 1)) * nw + jc + lr];
        }
    };
    __syncthreads();  // X_kk is read back as an operand below
    if (tk.a + 1 < nblk) load(tk.a + 1, tk.a);
    for (int ib = tk.a + 1; ib < nblk; ++ib) {
        const int ic = ib * NB, nbi = min(NB, nw - ic);
        acc_t acc[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{};
        for (int jb = tk.a; jb < ib; ++jb) {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                A[lr][lc + 4 * i] = lr < nbl ? ra[i] : T(0);
                B[lr][lc + 4 * i] = lc + 4 * i < nbk ? rb[i] : T(0);
            }
            __syncthreads();
            // the next product of this block row; X(jb + 1, kk) was written before this row
            if (jb + 1 < ib) load(ib, jb + 1);
            tile_product_mfma<T>(A, B, acc, row0, col0, l16, l4);
            __syncthreads();
        }
        // X_ik = -Linv_ii S: Linv_ii into A, the sum S into B
#pragma unroll
        for (int i = 0; i < PER; ++i)
            ra[i] = DB[(int64_t)(cg + c0 + ic + min(lc + 4 * i, nbi - 1)) * NB + min(lr, nbi - 1)];
#pragma unroll
        for (int i = 0; i < PER; ++i) A[lr][lc + 4 * i] = lr < nbi && lc + 4 * i < nbi ? ra[i] : T(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    B[row0 + 16 * a + Mfma<T>::row(l4, q)][col0 + 16 * b + l16] = acc[a][b][q];
        __syncthreads();
        acc_t out[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) out[a][b] = acc_t{};
        tile_product_mfma<T>(A, B, out, row0, col0, l16, l4);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = row0 + 16 * a + Mfma<T>::row(l4, q), c = col0 + 16 * b + l16;
                    if (r < nbi && c < nbk) X[(int64_t)(kc + c) * nw + ic + r] = -out[a][b][q];
                }
        __syncthreads();  // X_ik is visible to the next block row's loads; the LDS tiles are free
        if (ib + 1 < nblk) load(ib + 1, tk.a);
    }
}

// forward, wide steps: y_i = sum_{k <= i} X_ik t_k for 16 rows [ri0, ri0 + 16) of a step, one
// workgroup of 256 threads per (step, ri0): thread (row, column group g of 16) takes the
// 16-column batches g, g + 16, ... (a wave reads four 16-row column segments per load); the 16
// partial sums of a row are added in group order. t = the step's rows of y, less the children's
// update entries at the first step of a front (applied here to an LDS copy). The result goes
// to Z (the other workgroups still read y); k_copy_rows moves it to y. task: (step, ri0).
// LDS: SOLVE_STEP 8 + 16 16 8 entries.
template <typename T>
__device__ __attribute__((always_inline)) inline void fwd_inv_body(const DevSym& S, const SolveStep* __restrict__ steps,
                                                                   const Task tk, const T* __restrict__ XB,
                                                                   const T* __restrict__ Y, T* __restrict__ Z,
                                                                   const T* __restrict__ U, int nrhs) {
    T* t = reinterpret_cast<T*>(sweep_lds);  // step rows x 8
    T* red = t + SOLVE_STEP * 8;              // 16 groups x 16 rows x 8
    const SolveStep st = steps[tk.s];
    const int s = st.s, c0 = st.c0, c1 = st.c1, nw = c1 - c0;
    const int cg = S.super_ptr[s];
    const T* X = XB + st.xoff;
    const int ri0 = tk.a, nbi = min(16, nw - ri0), ncols = ri0 + nbi;
    const int tid = threadIdx.x, rl = tid & 15, g = tid >> 4;
    const int r = ri0 + min(rl, nbi - 1);
    for (int rb = 0; rb < nrhs; rb += 8) {
        const int nr = min(8, nrhs - rb);
        // t = the step's rows of y, less (at the front's first step) the children's update entries
        // of those rows, row by row from the pull map in child_list order
        const int32_t* pp = S.pm_ptr + S.pm_row[s];
        for (int q = tid; q < ncols * 8; q += 256) {
            const int i = q >> 3, rr = q & 7;
            T v = T(0);
            if (rr < nr) {
                v = Y[(int64_t)(cg + c0 + i) * nrhs + rb + rr];
                if (c0 == 0) {
                    const int e1 = pp[i + 1];
                    for (int e = pp[i]; e < e1; ++e) v -= U[(int64_t)S.pm_src[e] * nrhs + rb + rr];
                }
            }
            t[q] = v;
        }
        __syncthreads();
        T acc[8];
#pragma unroll
        for (int rr = 0; rr < 8; ++rr) acc[rr] = T(0);
        const T* xr = X + r;
        for (int cb = 16 * g; cb < ncols; cb += 256) {
            T l[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) l[jj] = xr[(int64_t)min(cb + jj, ncols - 1) * nw];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const T lv = cb + jj < ncols ? l[jj] : T(0);
                const T* tj = t + min(cb + jj, ncols - 1) * 8;
#pragma unroll
                for (int rr = 0; rr < 8; ++rr) acc[rr] += lv * tj[rr];
            }
        }
#pragma unroll
        for (int rr = 0; rr < 8; ++rr) red[(g * 16 + rl) * 8 + rr] = acc[rr];
        __syncthreads();
        if (tid < 16 * 8) {
            const int row = tid >> 3, rr = tid & 7;
            if (row < nbi && rr < nr) {
                T v = red[row * 8 + rr];
#pragma unroll
                for (int u = 1; u < 16; ++u) v += red[(u * 16 + row) * 8 + rr];
                Z[(int64_t)(cg + c0 + ri0 + row) * nrhs + rb + rr] = v;
            }
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fwd_inv(DevSym S, const SolveStep* __restrict__ steps,
                                                const Task* __restrict__ tasks, const T* __restrict__ XB,
                                                const T* __restrict__ Y, T* __restrict__ Z, const T* __restrict__ U,
                                                int nrhs) {
    fwd_inv_body<T>(S, steps, tasks[blockIdx.x], XB, Y, Z, U, nrhs);
}

// backward, wide steps: x_j = sum_{i >= j} X_ij t_i for 8 columns [col0, col0 + 8) of a step,
// one wave per (step, col0); t = the step's rows of y (k_sub_partials has subtracted the
// L21^T x partial sums), rows >= col0 staged in LDS. Lanes along the rows of X (coalesced); the 64 sums are
// reduce-scattered across the lanes. t (y_step less the partial sums) comes from Z, where
// k_sub_partials formed it, and x goes to y directly (no task of the launch reads those rows of y).
// task: (step, col0). LDS: SOLVE_STEP 8 entries.
template <typename T>
__global__ void __launch_bounds__(64) k_bwd_inv(DevSym S, const SolveStep* __restrict__ steps,
                                               const Task* __restrict__ tasks, const T* __restrict__ XB,
                                               T* __restrict__ Y, const T* __restrict__ Z, int nrhs) {
    T* t = reinterpret_cast<T*>(sweep_lds);  // rows [col0, nw) of the step x 8
    const Task tk = tasks[blockIdx.x];
    const SolveStep st = steps[tk.s];
    const int s = st.s, c0 = st.c0, c1 = st.c1, nw = c1 - c0;
    const int cg = S.super_ptr[s];
    const T* X = XB + st.xoff;
    const int col0 = tk.a, cend = min(nw, col0 + 8);
    const int lane = threadIdx.x;
    for (int rb = 0; rb < nrhs; rb += 8) {
        const int nr = min(8, nrhs - rb);
        for (int q = lane; q < (nw - col0) * 8; q += 64) {
            const int i = col0 + (q >> 3), rr = q & 7;
            t[q] = rr < nr ? Z[(int64_t)(cg + c0 + i) * nrhs + rb + rr] : T(0);
        }
        __syncthreads();
        T a[64];
#pragma unroll
        for (int v = 0; v < 64; ++v) a[v] = T(0);
#pragma unroll 2
        for (int i = col0 + lane; i < nw; i += 64) {
            const T* ti = t + (i - col0) * 8;
            T x[8];
#pragma unroll
            for (int rr = 0; rr < 8; ++rr) x[rr] = ti[rr];
            const T* pc = X + (int64_t)col0 * nw + i;  // column col0 + c; columns past cend read cend - 1
            T l[8];  // the row's eight loads issued together (the FP32 build otherwise waits per load)
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                l[c] = *pc;
                pc += col0 + c + 1 < cend ? nw : 0;
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int rr = 0; rr < 8; ++rr) a[c * 8 + rr] += l[c] * x[rr];
        }
        const T v = wave_reduce_scatter64<T>(a);
        const int col = col0 + (lane >> 3), rr = lane & 7;
        if (col < cend && rr < nr) Y[(int64_t)(cg + c0 + col) * nrhs + rb + rr] = v;
        __syncthreads();
    }
}

// rows [r0, r0 + nr) of a wide step: Z = y less the L21^T x partial sums of those rows, in chunk
// order (the input of k_bwd_inv). task: (step, r0, rows)
template <typename T>
__device__ __attribute__((always_inline)) inline void sub_partials_body(const DevSym& S,
                                                                        const SolveStep* __restrict__ steps,
                                                                        const Task tk, const T* __restrict__ Y,
                                                                        T* __restrict__ Z, const T* __restrict__ part,
                                                                        int nrhs) {
    const SolveStep st = steps[tk.s];
    const int nw = st.c1 - st.c0;
    const int64_t o = (int64_t)(S.super_ptr[st.s] + st.c0 + tk.a) * nrhs;
    const T* pp = part + (st.poff + tk.a) * nrhs;
    for (int q = threadIdx.x; q < tk.b * nrhs; q += 256) {
        T v = Y[o + q];
        for (int ch = 0; ch < st.nch; ++ch) v -= pp[(int64_t)ch * nw * nrhs + q];
        Z[o + q] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_sub_partials(DevSym S, const SolveStep* __restrict__ steps,
                                                     const Task* __restrict__ tasks, const T* __restrict__ Y,
                                                     T* __restrict__ Z, const T* __restrict__ part, int nrhs) {
    sub_partials_body<T>(S, steps, tasks[blockIdx.x], Y, Z, part, nrhs);
}

// y rows [r0, r0 + nr) of a wide step from Z (the result of k_fwd_inv / k_bwd_inv).
// task: (step, r0, rows)
template <typename T>
__global__ void __launch_bounds__(256) k_copy_rows(DevSym S, const SolveStep* __restrict__ steps,
                                                  const Task* __restrict__ tasks, T* __restrict__ Y,
                                                  const T* __restrict__ Z, int nrhs) {
    const Task tk = tasks[blockIdx.x];
    const SolveStep st = steps[tk.s];
    const int64_t o = (int64_t)(S.super_ptr[st.s] + st.c0 + tk.a) * nrhs;
    for (int q = threadIdx.x; q < tk.b * nrhs; q += 256) Y[o + q] = Z[o + q];
}

// ---------------------------------------------------------------------------

// One launch for the two independent halves of a step group (different fronts, disjoint rows):
// forward, the narrow steps (the first ndiag workgroups, tasks of k_fwd_diag) and the wide steps'
// inverse products (tasks of k_fwd_inv); backward, the narrow steps (k_bwd_diag) and the wide
// steps' t = y less the partial sums (k_sub_partials). Each workgroup runs its kernel's body.
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_group(DevSym S, const SolveStep* __restrict__ steps,
                                                  const Task* __restrict__ tdiag, int ndiag,
                                                  const Task* __restrict__ tinv, const T* __restrict__ factor,
                                                  const T* __restrict__ DB, const T* __restrict__ XB,
                                                  T* __restrict__ Y, T* __restrict__ Z, T* __restrict__ U, int nrhs) {
    if ((int)blockIdx.x < ndiag)
        fwd_diag_body<T>(S, steps, tdiag[blockIdx.x], factor, DB, Y, U, nrhs);
    else
        fwd_inv_body<T>(S, steps, tinv[blockIdx.x - ndiag], XB, Y, Z, U, nrhs);
}

template <typename T>
__global__ void __launch_bounds__(256) k_bwd_group(DevSym S, const SolveStep* __restrict__ steps,
                                                  const Task* __restrict__ tdiag, int ndiag,
                                                  const Task* __restrict__ tsub, const T* __restrict__ factor,
                                                  const T* __restrict__ DB, T* __restrict__ Y, T* __restrict__ Z,
                                                  const T* __restrict__ part, int nrhs) {
    if ((int)blockIdx.x < ndiag)
        bwd_diag_body<T>(S, steps, tdiag[blockIdx.x], factor, DB, Y, part, nrhs);
    else
        sub_partials_body<T>(S, steps, tsub[blockIdx.x - ndiag], Y, Z, part, nrhs);
}

// ---------------------------------------------------------------------------

template <typename T>
void launch_block_inverse(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* DB, int bs) {
    if (bs == 16)
        hipLaunchKernelGGL((k_block_inverse<T, 16>), dim3((n + 3) / 4), dim3(64), 0, st, S, tasks, n, factor, DB);
    else if (bs == 32)
        hipLaunchKernelGGL((k_block_inverse<T, 32>), dim3((n + 1) / 2), dim3(64), 0, st, S, tasks, n, factor, DB);
    else
        hipLaunchKernelGGL((k_block_inverse<T, NB>), dim3(n), dim3(64), 0, st, S, tasks, n, factor, DB);
}

template <typename T>
void launch_fwd_diag(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, int threads,
                     const T* factor, const T* DB, T* Y, T* U, int nrhs) {
    const size_t lds = 2 * NB * 8 * sizeof(T);
    hipLaunchKernelGGL(k_fwd_diag<T>, dim3(n), dim3(threads), lds, st, S, steps, tasks, factor, DB, Y, U, nrhs);
}

template <typename T>
void launch_fwd_group(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tdiag, int ndiag,
                      const Task* tinv, int ninv, const T* factor, const T* DB, const T* XB, T* Y, T* Z, T* U,
                      int nrhs) {
    const size_t lds = std::max(2 * NB * 8, SOLVE_STEP * 8 + 16 * 16 * 8) * sizeof(T);
    hipLaunchKernelGGL(k_fwd_group<T>, dim3(ndiag + ninv), dim3(256), lds, st, S, steps, tdiag, ndiag, tinv, factor,
                       DB, XB, Y, Z, U, nrhs);
}

template <typename T>
void launch_bwd_group(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tdiag, int ndiag,
                      const Task* tsub, int nsub, const T* factor, const T* DB, T* Y, T* Z, const T* part,
                      int nrhs) {
    const size_t lds = NB * 8 * sizeof(T);
    hipLaunchKernelGGL(k_bwd_group<T>, dim3(ndiag + nsub), dim3(256), lds, st, S, steps, tdiag, ndiag, tsub, factor,
                       DB, Y, Z, part, nrhs);
}

template <typename T>
void launch_fwd_rows(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                     const T* factor, T* Y, const T* Z, T* U, int nrhs) {
    const size_t lds = (SOLVE_STEP * 8 + 8 * 64 * 8) * sizeof(T);
    hipLaunchKernelGGL(k_fwd_rows<T>, dim3(n), dim3(512), lds, st, S, steps, tasks, factor, Y, Z, U, nrhs);
}

template <typename T>
void launch_bwd_l21t(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                     const T* factor, const T* Y, T* part, int nrhs) {
    hipLaunchKernelGGL(k_bwd_l21t<T>, dim3(n), dim3(64), 0, st, S, steps, tasks, factor, Y, part, nrhs);
}

template <typename T>
void launch_bwd_diag(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, int threads,
                     const T* factor, const T* DB, T* Y, const T* part, int nrhs) {
    const size_t lds = NB * 8 * sizeof(T);
    hipLaunchKernelGGL(k_bwd_diag<T>, dim3(n), dim3(threads), lds, st, S, steps, tasks, factor, DB, Y, part, nrhs);
}

template <typename T>
void launch_step_inverse(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                         const T* factor, const T* DB, T* XB) {
    hipLaunchKernelGGL(k_step_inverse<T>, dim3(n), dim3(256), 0, st, S, steps, tasks, factor, DB, XB);
}

template <typename T>
void launch_fwd_inv(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, const T* XB,
                    const T* Y, T* Z, const T* U, int nrhs) {
    const size_t lds = (SOLVE_STEP * 8 + 16 * 16 * 8) * sizeof(T);
    hipLaunchKernelGGL(k_fwd_inv<T>, dim3(n), dim3(256), lds, st, S, steps, tasks, XB, Y, Z, U, nrhs);
}

template <typename T>
void launch_bwd_inv(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, const T* XB,
                    T* Y, const T* Z, int nrhs) {
    const size_t lds = SOLVE_STEP * 8 * sizeof(T);
    hipLaunchKernelGGL(k_bwd_inv<T>, dim3(n), dim3(64), lds, st, S, steps, tasks, XB, Y, Z, nrhs);
}

template <typename T>
void launch_sub_partials(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n,
                         const T* Y, T* Z, const T* part, int nrhs) {
    hipLaunchKernelGGL(k_sub_partials<T>, dim3(n), dim3(256), 0, st, S, steps, tasks, Y, Z, part, nrhs);
}

template <typename T>
void launch_copy_rows(hipStream_t st, const DevSym& S, const SolveStep* steps, const Task* tasks, int n, T* Y,
                      const T* Z, int nrhs) {
    hipLaunchKernelGGL(k_copy_rows<T>, dim3(n), dim3(256), 0, st, S, steps, tasks, Y, Z, nrhs);
}

#define INST(T)                                                                                                     \
    template void launch_block_inverse<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, int);                 \
    template void launch_fwd_diag<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, int, const T*, \
                                     const T*, T*, T*, int);                                                       \
    template void launch_fwd_rows<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, const T*, T*, \
                                     const T*, T*, int);                                                           \
    template void launch_bwd_l21t<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, const T*,     \
                                     const T*, T*, int);                                                           \
    template void launch_bwd_diag<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, int, const T*, \
                                     const T*, T*, const T*, int);                                                \
    template void launch_step_inverse<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int,          \
                                         const T*, const T*, T*);                                                  \
    template void launch_fwd_inv<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, const T*,      \
                                    const T*, T*, const T*, int);                                                  \
    template void launch_bwd_inv<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, const T*,      \
                                    T*, const T*, int);                                                            \
    template void launch_sub_partials<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int,          \
                                         const T*, T*, const T*, int);                                             \
    template void launch_copy_rows<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, T*,         \
                                      const T*, int);                                                              \
    template void launch_fwd_group<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, const Task*, \
                                      int, const T*, const T*, const T*, T*, T*, T*, int);                         \
    template void launch_bwd_group<T>(hipStream_t, const DevSym&, const SolveStep*, const Task*, int, const Task*, \
                                      int, const T*, const T*, T*, T*, const T*, int);
INST(double)
INST(float)
#undef INST

}  // namespace sds
