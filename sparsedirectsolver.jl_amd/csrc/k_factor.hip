// Numeric factorization kernels (gfx950). Each front writes only its own panel
// and contribution block; the extend-add is owner-pull (children in child_list
// order) and never uses atomics, so a factorization is bitwise reproducible.
//
// Small fronts (the whole f x f front fits the LDS of one workgroup) are
// assembled and factored by one fused kernel, one workgroup per front, all the
// small fronts of a tree level in one launch. Large fronts are assembled in
// global memory (panel in the factor buffer, contribution block in the CB pool)
// by column-chunk tasks, then factored block column by block column (NB wide):
// diagonal block in LDS, row-tile triangular solves, and Schur updates as 64x64
// tiles on the matrix cores; the contribution block gets one rank-w update with
// K = w after the panel is done.
//
// Reference counterparts: regime B fused kernel src/numeric/front.jl:154-206,
// assembly src/numeric/assembly.jl:18-252, regime C panel src/numeric/factorize.jl:240-269,
// statistics src/numeric/factorize.jl:13-48.
#include "k_dense.h"
#include "k_front.h"
#include "sds_common.h"
#include "sds_kernels.h"

// phase marks of the large-front step chain; empty unless a timing build defines them
#ifndef SDS_STEP_MARK
#define SDS_STEP_MARK(i)
#endif
#ifndef SDS_POTRF_MARK
#define SDS_POTRF_MARK(i)
#endif

namespace sds {

// 1/sqrt(a), a > 0: the hardware estimate refined by two Newton steps (the pivots of the
// large fronts' diagonal blocks; the result is within an ulp or two, well inside the
// factor's accuracy)
__device__ inline double rsqrt_nr(double a) {
    double y = __builtin_amdgcn_rsq(a);
    const double h = 0.5 * a;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// ---------------------------------------------------------------------------
// small fronts: fused assembly + dense Cholesky of the whole front in LDS

template <typename T>
__global__ void __launch_bounds__(256) k_front_small_chol(DevSym S, const int32_t* __restrict__ fronts,
                                                         const double* __restrict__ nz, T* __restrict__ factor,
                                                         T* __restrict__ cbpool, T* __restrict__ DB,
                                                         int32_t* __restrict__ status, double eps) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    __shared__ int sh;
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int tid = threadIdx.x;
    assemble_dense<T>(S, s, nz, cbpool, F, f, 0);
    // blocked Cholesky of the w pivot columns over the whole front (f <= 128)
    const int st = front_chol_lds<T, 2>(F, f, w, eps, &sh);
    // the panel (strict upper triangle zero) and the contribution block's lower triangle, a wave
    // per column, lanes along the rows
    const int lane = tid & 63, wv = tid >> 6;
    T* P = factor + S.panel_off[s];
    for (int j = wv; j < w; j += 4)
        for (int i = lane; i < f; i += 64) P[(int64_t)j * f + i] = i >= j ? F[j * f + i] : T(0);
    const int64_t co = S.cb_off[s];
    if (co >= 0) {
        T* C = cbpool + co;
        for (int jj = wv; jj < m; jj += 4)
            for (int ii = jj + lane; ii < m; ii += 64) C[(int64_t)jj * m + ii] = F[(w + jj) * f + w + ii];
    }
    // the diagonal block's inverse for the sweeps (fronts of at most FUSED_INV_MAX columns)
    if (w <= FUSED_INV_MAX) diag_inverse_lds<T>(F, f, w, DB + (int64_t)S.super_ptr[s] * NB);
    if (tid == 0) status[s] = st;
}

// The same small fronts factored column by column: per pivot column two barriers, the column
// scaled by 1/d, then the whole trailing lower triangle updated by all threads (the reference's
// regime-B form, src/numeric/front.jl:154-206). The entries receive exactly the operations of
// k_front_small_chol in the same order, so both give the same factor bit for bit; which one is
// faster depends on the fronts (the autotuner chooses).
template <typename T>
__global__ void __launch_bounds__(256) k_front_col_chol(DevSym S, const int32_t* __restrict__ fronts,
                                                       const double* __restrict__ nz, T* __restrict__ factor,
                                                       T* __restrict__ cbpool, T* __restrict__ DB,
                                                       int32_t* __restrict__ status, double eps) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    __shared__ int32_t st;
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) st = 0;
    assemble_dense<T>(S, s, nz, cbpool, F, f, 0);
    for (int j = 0; j < w; ++j) {
        if (tid == 0) {
            const double a = (double)F[j * f + j];
            if (chol_accept(a, eps))
                F[j * f + j] = T(sqrt(fmax(a, eps)));
            else
                st = j + 1;
        }
        __syncthreads();
        if (st) break;
        const T rd = T(1) / F[j * f + j];
        for (int i = j + 1 + tid; i < f; i += nt) F[j * f + i] = F[j * f + i] * rd;
        __syncthreads();
        // the lower triangle of the trailing square: one wave per column, lanes along the rows
        const int r = f - j - 1;
        const T* lj = F + j * f + j + 1;
        for (int kk = tid >> 6; kk < r; kk += nt >> 6) {
            const T lk = lj[kk];
            T* col = F + (j + 1 + kk) * f + j + 1;
            for (int ii = kk + (tid & 63); ii < r; ii += 64) {
                const T prod = lj[ii] * lk;
                col[ii] = col[ii] - prod;
            }
        }
        __syncthreads();
    }
    const int lane = tid & 63, wv = tid >> 6;
    T* P = factor + S.panel_off[s];
    for (int j = wv; j < w; j += 4)
        for (int i = lane; i < f; i += 64) P[(int64_t)j * f + i] = i >= j ? F[j * f + i] : T(0);
    const int64_t co = S.cb_off[s];
    if (co >= 0) {
        T* C = cbpool + co;
        for (int jj = wv; jj < m; jj += 4)
            for (int ii = jj + lane; ii < m; ii += 64) C[(int64_t)jj * m + ii] = F[(w + jj) * f + w + ii];
    }
    if (w <= FUSED_INV_MAX) diag_inverse_lds<T>(F, f, w, DB + (int64_t)S.super_ptr[s] * NB);
    if (tid == 0) status[s] = st;
}

// Regime A: one workgroup per leaf subtree, walking its fronts in postorder (sub_nodes) with the
// multifrontal stack in LDS (src/numeric/subtree.jl:274-316 is the reference's form over packed
// triangles; here every front is a dense f x f square and every contribution block an m x m
// square). A front at lfront[v] sits just above its children's blocks: zeroed, A added (one
// thread per run of entries with one destination, in nzval order), the children's blocks added
// in child_list order (the order of assemble_dense, so the front is bitwise the same as in the
// level kernels), factored by front_chol_lds; its panel and diagonal-block inverse go to global
// memory. The subtree root's block goes to the contribution-block pool for its parent; any other
// front's block moves down to lcb[v], where its first child's block began: every destination lies
// below its source and the gap grows along the copy, so a chunk read whole before it is written
// never overwrites a source still to be read.
template <typename T>
__global__ void __launch_bounds__(256) k_subtree_chol(DevSym S, const int32_t* __restrict__ sub_ptr,
                                                     const int32_t* __restrict__ sub_nodes, int sub0,
                                                     const int32_t* __restrict__ lfront, const int32_t* __restrict__ lcb,
                                                     const double* __restrict__ nz, T* __restrict__ factor,
                                                     T* __restrict__ cbpool, T* __restrict__ DB,
                                                     int32_t* __restrict__ status, double eps) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* stack = reinterpret_cast<T*>(smem_raw);
    __shared__ int sh;
    const int t = sub0 + blockIdx.x;
    const int n0 = sub_ptr[t], n1 = sub_ptr[t + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int kn = n0; kn < n1; ++kn) {
        const int v = sub_nodes[kn];
        const int f = S.fnrows[v], w = S.fncols[v], m = f - w;
        T* F = stack + lfront[v];
        for (int q = tid; q < f * f; q += 256) F[q] = T(0);
        __syncthreads();
        const int64_t a0 = S.asm_ptr[v], a1 = S.asm_ptr[v + 1];
        for (int64_t k = a0 + tid; k < a1; k += 256) {
            const int64_t d = S.asm_dst[k];
            if (k > a0 && S.asm_dst[k - 1] == d) continue;
            double val = 0.0;
            for (int64_t q = k; q < a1 && S.asm_dst[q] == d; ++q) val += nz[S.asm_src[q]];
            F[d >= 0 ? d : -(d + 1)] += T(val);
        }
        __syncthreads();
        for (int k = S.child_ptr[v]; k < S.child_ptr[v + 1]; ++k) {
            const int c = S.child_list[k];
            const int mc = S.fnrows[c] - S.fncols[c];
            const T* C = stack + lcb[c];
            const int32_t* ri = S.relind + S.relind_ptr[c];
            for (int q = tid; q < mc * mc; q += 256) {
                const int jj = q / mc, ii = q - jj * mc;
                if (ii >= jj) F[ri[jj] * f + ri[ii]] += C[q];
            }
            __syncthreads();
        }
        const int st = front_chol_lds<T, 2>(F, f, w, eps, &sh);
        T* P = factor + S.panel_off[v];
        for (int j = wv; j < w; j += 4)
            for (int i = lane; i < f; i += 64) P[(int64_t)j * f + i] = i >= j ? F[j * f + i] : T(0);
        if (w <= FUSED_INV_MAX) diag_inverse_lds<T>(F, f, w, DB + (int64_t)S.super_ptr[v] * NB);
        if (kn == n1 - 1) {
            const int64_t co = S.cb_off[v];
            if (co >= 0) {
                T* C = cbpool + co;
                for (int jj = wv; jj < m; jj += 4)
                    for (int ii = jj + lane; ii < m; ii += 64) C[(int64_t)jj * m + ii] = F[(w + jj) * f + w + ii];
            }
        } else {
            T* D = stack + lcb[v];
            for (int q0 = 0; q0 < m * m; q0 += 1024) {
                T val[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = q0 + tid + 256 * u;
                    const int jj = q / max(m, 1), ii = q - jj * m;
                    val[u] = q < m * m ? F[(w + jj) * f + w + ii] : T(0);
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = q0 + tid + 256 * u;
                    if (q < m * m) D[q] = val[u];
                }
                __syncthreads();
            }
        }
        if (tid == 0) status[v] = st;
        __syncthreads();
    }
}

// fronts of at most TINY_F rows: one wave per front, four per workgroup (no block barriers). The
// wave assembles the front in its own LDS square (zero, A, then the children in child_list
// order, as assemble_dense), then each lane takes one row into registers and the wave factors
// the w pivot columns one by one: the pivot broadcast from its lane, the rows below scaled by
// 1/d, every later column k of row i given -= L(i, j) L(k, j) with L(k, j) read from lane k. An
// entry's updates come in column order as separate products and differences, the arithmetic of
// the blocked kernel. The panel and contribution block are written from the registers, and for
// fronts of at most FUSED_INV_MAX columns the diagonal block's inverse (column c by lane c, the
// arithmetic of the per-block inverse kernel).
template <typename T>
__global__ void __launch_bounds__(256) k_front_tiny_chol(DevSym S, const int32_t* __restrict__ fronts, int nfronts,
                                                        const double* __restrict__ nz, T* __restrict__ factor,
                                                        T* __restrict__ cbpool, T* __restrict__ DB,
                                                        int32_t* __restrict__ status, double eps) {
    __shared__ T Fs[4][TINY_F * TINY_F];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int id = blockIdx.x * 4 + wv;
    if (id >= nfronts) return;  // the whole wave: no block barriers below
    T* F = Fs[wv];
    const int s = fronts[id];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    for (int q = lane; q < f * f; q += 64) F[q] = T(0);
    wave_sync();
    // A: one lane sums each run of entries with one destination, in nzval order (distinct runs
    // have distinct destinations)
    const int64_t a0 = S.asm_ptr[s], a1 = S.asm_ptr[s + 1];
    for (int64_t k = a0 + lane; k < a1; k += 64) {
        const int64_t d = S.asm_dst[k];
        if (k > a0 && S.asm_dst[k - 1] == d) continue;
        double v = 0.0;
        for (int64_t q = k; q < a1 && S.asm_dst[q] == d; ++q) v += nz[S.asm_src[q]];
        F[d >= 0 ? d : -(d + 1)] += T(v);
    }
    wave_sync();
    for (int k = S.child_ptr[s]; k < S.child_ptr[s + 1]; ++k) {
        const int c = S.child_list[k];
        const int64_t co = S.cb_off[c];
        if (co < 0) continue;
        const int mc = S.fnrows[c] - S.fncols[c];
        const T* C = cbpool + co;
        const int32_t* ri = S.relind + S.relind_ptr[c];
        const int rl = lane < mc ? ri[lane] : 0;
        for (int jj = 0; jj < mc; ++jj) {
            const int pj = __builtin_amdgcn_readlane(rl, jj);
            if (lane >= jj && lane < mc) F[pj * f + rl] += C[(int64_t)jj * mc + lane];
        }
        wave_sync();
    }
    // row `lane` into registers
    T r[TINY_F];
#pragma unroll
    for (int k = 0; k < TINY_F; ++k) r[k] = (k < f && lane < f) ? F[k * f + lane] : T(0);
    int st = 0;
#pragma unroll
    for (int j = 0; j < TINY_F; ++j) {
        if (j < w && !st) {
            const double a = (double)lane_value(r[j], j);
            if (!chol_accept(a, eps)) {
                st = j + 1;
            } else {
                const T d = T(sqrt(fmax(a, eps)));
                const T rd = T(1) / d;
                if (lane == j) r[j] = d;
                else if (lane > j) r[j] = r[j] * rd;
#pragma unroll
                for (int k = j + 1; k < TINY_F; ++k) {
                    if (k < f) {
                        const T lk = lane_value(r[j], k);
                        if (lane >= k) {
                            const T prod = r[j] * lk;
                            r[k] = r[k] - prod;
                        }
                    }
                }
            }
        }
    }
    T* P = factor + S.panel_off[s];
#pragma unroll
    for (int j = 0; j < TINY_F; ++j)
        if (j < w && lane < f) P[(int64_t)j * f + lane] = lane >= j ? r[j] : T(0);
    const int64_t co = S.cb_off[s];
    if (co >= 0) {
        T* C = cbpool + co;
#pragma unroll
        for (int k = 0; k < TINY_F; ++k)
            if (k >= w && k < f && lane >= k && lane < f) C[(int64_t)(k - w) * m + lane - w] = r[k];
    }
    if (w <= FUSED_INV_MAX) {
        // column `lane` of L11^-1: L(rr, mm) is row rr's register mm
        T x[FUSED_INV_MAX];
#pragma unroll
        for (int q = 0; q < FUSED_INV_MAX; ++q) x[q] = q == lane ? T(1) : T(0);
#pragma unroll
        for (int mm = 0; mm < FUSED_INV_MAX; ++mm) {
            if (mm < w) {
                const T rinv = T(1) / lane_value(r[mm], mm);
                x[mm] = x[mm] * rinv;
                const T xm = x[mm];
#pragma unroll
                for (int rr = mm + 1; rr < FUSED_INV_MAX; ++rr)
                    if (rr < w) x[rr] = fma(-lane_value(r[mm], rr), xm, x[rr]);
            }
        }
        if (lane < w) {
            T* o = DB + (int64_t)(S.super_ptr[s] + lane) * NB;
#pragma unroll
            for (int q = 0; q < FUSED_INV_MAX; ++q) o[q] = q < w ? x[q] : T(0);
#pragma unroll
            for (int q = FUSED_INV_MAX; q < NB; ++q) o[q] = T(0);
        }
    }
    if (lane == 0) status[s] = st;
}

// ---------------------------------------------------------------------------
// large fronts

// address of front entry (i, j), i >= j: panel column or contribution block
template <typename T>
__device__ inline T* front_at(T* P, T* C, int f, int w, int i, int j) {
    return j < w ? P + (int64_t)j * f + i : C + (int64_t)(j - w) * (f - w) + (i - w);
}

// zero a chunk of columns of a large front and add A into its panel columns
// task: (s, col0, ncols)
template <typename T>
__global__ void __launch_bounds__(256) k_big_zero_scatter(DevSym S, const Task* __restrict__ tasks,
                                                         const double* __restrict__ nz, T* __restrict__ factor,
                                                         T* __restrict__ cbpool) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], m = f - w;
    T* P = factor + S.panel_off[s];
    T* C = cbpool + (S.cb_off[s] >= 0 ? S.cb_off[s] : 0);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int j1 = min(t.a + t.b, f);
    for (int j = t.a; j < j1; ++j) {
        if (j < w) {
            for (int i = tid; i < f; i += nt) P[(int64_t)j * f + i] = T(0);
        } else {  // the lower triangle: the only part of a contribution block that is read
            for (int i = (j - w) + tid; i < m; i += nt) C[(int64_t)(j - w) * m + i] = T(0);
        }
    }
    __syncthreads();
    const int c0 = S.super_ptr[s];
    for (int j = t.a; j < min(j1, w); ++j) {
        const int64_t a0 = S.acol_ptr[c0 + j], a1 = S.acol_ptr[c0 + j + 1];
        for (int64_t k = a0 + tid; k < a1; k += nt) {
            const int64_t d = S.asm_dst[k];
            if (k > a0 && S.asm_dst[k - 1] == d) continue;
            double v = 0.0;
            for (int64_t q = k; q < a1 && S.asm_dst[q] == d; ++q) v += nz[S.asm_src[q]];
            const int64_t loc = d >= 0 ? d : -(d + 1);
            P[loc] += T(v);
        }
    }
}

// extend-add, owner-pull: task (s, pc0, e0, ne) owns the parent entries of the columns
// [pc0, pc1) (at most EA_COLS, never across w) in one range of rows, and adds into them the
// entries of the children that land there, child by child in child_list order (a barrier
// between children: two children can hit one parent entry). Its ne entries, tasks[e0 ..] from
// the launch's first record: (child c, first child column jj0 landing in the block, child rows
// [ii0, ii1)). A thread takes one child row at a time (its parent row read once) and the
// block's columns of that row: the child reads and the parent read-modify-writes are coalesced
// along the rows, several in flight per thread. Each parent entry gets its children's entries
// in child_list order, as in the reference's owner-pull extend-add (src/numeric/assembly.jl:
// 180-198), so the sums are bitwise those of a child-by-child sweep.
template <typename T>
__global__ void __launch_bounds__(256) k_big_extend_add(DevSym S, const Task* __restrict__ tasks,
                                                       T* __restrict__ factor, T* __restrict__ cbpool) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int pc0 = t.a, pc1 = min(pc0 + EA_COLS, pc0 < w ? w : f);
    T* P = factor + S.panel_off[s];
    T* Cs = cbpool + (S.cb_off[s] >= 0 ? S.cb_off[s] : 0);
    for (int e = 0; e < t.c; ++e) {
        const Task E = tasks[t.b + e];
        const int c = E.s, jj0 = E.a, mc = S.fnrows[c] - S.fncols[c];
        const T* Cc = cbpool + S.cb_off[c];
        const int32_t* ri = S.relind + S.relind_ptr[c];
        int pj[EA_COLS];
#pragma unroll
        for (int q = 0; q < EA_COLS; ++q) pj[q] = jj0 + q < mc ? ri[jj0 + q] : f;  // f: past the block
        for (int ii = E.b + threadIdx.x; ii < E.c; ii += 256) {
            const int pi = ri[ii];
            T v[EA_COLS], old[EA_COLS];
            T* dst[EA_COLS];
#pragma unroll
            for (int q = 0; q < EA_COLS; ++q) {
                const int jj = jj0 + q;
                const bool in = pj[q] < pc1 && ii >= jj;
                // parent entry (pi, pj): panel column pj < w, else the contribution block
                dst[q] = !in ? nullptr
                             : pj[q] < w ? P + (int64_t)pj[q] * f + pi : Cs + (int64_t)(pj[q] - w) * m + (pi - w);
                // the child's block is read once: keep it out of the caches the parent's entries use
                v[q] = in ? __builtin_nontemporal_load(Cc + (int64_t)jj * mc + ii) : T(0);
            }
#pragma unroll
            for (int q = 0; q < EA_COLS; ++q) old[q] = dst[q] ? *dst[q] : T(0);
#pragma unroll
            for (int q = 0; q < EA_COLS; ++q)
                if (dst[q]) *dst[q] = old[q] + v[q];
        }
        __syncthreads();  // the next child may add into the same parent entries
    }
}

// One block row p of X = L^-1, blocked by 16, for the 64 x 64 factor in D (D[j][i] = L(i, j);
// rd the pivots' reciprocals), into X (X[j][i] = X(i, j)), by one wave: the diagonal block
// X_pp by forward substitution (lane c: column c, sixteen steps), written when this wave also
// forms a piece or diag is set; piece j < p: X_pj = -X_pp sum_{k=j}^{p-1} L_pk X_kj with 16 x 16
// MFMA products (the second product's K runs through the first's accumulator registers, so T
// needs no staging). X_kj, k < p, must be in X already.
template <typename T>
__device__ __attribute__((always_inline)) inline void inverse_row(const T (*D)[NB + 1], const T* rd, T (*X)[NB + 1],
                                                                  int p, int j, bool diag) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x & 63, l16 = lane & 15, l4 = lane >> 4;
    const int b = 16 * p;
    {
        const int c = l16;
        T x[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) x[r] = r == c ? T(1) : T(0);
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            x[m] = x[m] * rd[b + m];
            const T xm = x[m];
#pragma unroll
            for (int r = m + 1; r < 16; ++r) x[r] = fma(-D[b + m][b + r], xm, x[r]);
        }
        if (lane < 16 && (diag || j < p))
#pragma unroll
            for (int r = 0; r < 16; ++r) X[b + c][b + r] = x[r];
    }
    if (j >= p) return;
    acc_t t = acc_t{};
    for (int k = j; k < p; ++k)
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
            t = Mfma<T>::mac(D[16 * k + 4 * s4 + l4][b + l16], X[16 * j + l16][16 * k + 4 * s4 + l4], t);
    acc_t x2 = acc_t{};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int kx = Mfma<T>::row(l4, q);  // t's element q of lane l is T(kx, l % 16)
        x2 = Mfma<T>::mac(-X[b + kx][b + l16], t[q], x2);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) X[16 * j + l16][b + Mfma<T>::row(l4, q)] = x2[q];
}

// Cholesky of the 64 x 64 block D in LDS (D[j][i] = L(i, j), lower part; padded with the
// identity past nb) and its inverse X = L^-1 (X[j][i] = X(i, j); its blocks above the diagonal
// blocks are left unwritten), four waves: per 16-column panel, wave 0 factors the panel's rows
// in registers (lane i holds row kb + i; the pivot and the next column's entry move between
// lanes by v_readlane, each finished column goes to D and the lanes read its further entries
// back as LDS broadcasts; 1/sqrt is
// the hardware estimate with two Newton steps) while waves 1-3 form the previous panel's block
// row of X; then the four waves apply the panel to the trailing lower 16 x 16 tiles on the
// matrix cores (K = 16). The last block row of X follows the loop. The upper triangle of D
// holds scratch values afterwards. rd gets the pivots' reciprocals (1 past nb); a failed pivot
// (1-based within the block) is left in fail, and X is then incomplete. Ends with a barrier.
struct NoRound0 {
    __device__ void operator()(int) const {}
};

// (pre: work for waves 1-3 during the first panel, which only reads and writes D's first 16
// columns; it may write D's other columns, and read X, which is not written before the second
// panel)
template <typename T, typename Pre = NoRound0>
__device__ __attribute__((always_inline)) inline void block_potrf_inv(T (*D)[NB + 1], T* rd, int* fail, int nb,
                                                                      double eps, T (*X)[NB + 1],
                                                                      Pre pre = Pre()) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int tid = threadIdx.x, wv = tid >> 6;
    if (tid == 0) *fail = 0;
    __syncthreads();
    int p = 0;  // the panel of this round
    for (int kb = 0; kb < nb; kb += 16, ++p) {
        // the lane index is re-read each panel so that the per-column lane masks are formed
        // where they are used instead of being kept (and spilled) across the loop
        int ii = tid & (NB - 1);
        asm volatile("" : "+v"(ii));
        if (wv == 0) {
            const int row = kb + ii, rowc = min(row, NB - 1);
            T p16[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const T v = D[kb + c][rowc];
                p16[c] = row < NB ? v : T(0);
            }
            T myr = T(1);         // lane c keeps the reciprocal of pivot c
            double mypiv = 1.0;   // and pivot c, for the acceptance test after the panel
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const double piv = (double)lane_bcast(p16[c], c);
                const double pe = fmax(piv, eps);
                const T r = T(rsqrt_nr(pe));
                // rows below the pivot get L(i, c) = a(i, c) r, the pivot row sqrt(max(a, eps)); the
                // rows above (the upper triangle) only hold scratch values
                const bool isc = ii == c;
                p16[c] = (isc ? T(pe) : p16[c]) * r;
                myr = isc ? r : myr;
                mypiv = isc ? piv : mypiv;
                // column c is final: it goes to its place in D. The entry L(kb + c + 1, kb + c)
                // that the next pivot waits for moves by v_readlane; the lanes read the entries
                // further down, needed only by later columns, back from D as LDS broadcasts (two
                // per read, their latency under the next pivot's; one wave's LDS accesses
                // complete in order)
                if (row < NB) D[kb + c][row] = p16[c];
                if (c + 1 < 16) p16[c + 1] = fma(-p16[c], lane_bcast(p16[c], c + 1), p16[c + 1]);
#pragma unroll
                for (int cc = c + 2; cc < 16; ++cc) p16[cc] = fma(-p16[c], D[kb + c][kb + cc], p16[cc]);
            }
            if (ii < 16) rd[kb + ii] = myr;
            // the first pivot of the panel that fails the acceptance test (1-based), if any
            const unsigned long long bad = __ballot(ii < 16 && kb + ii < nb && !chol_accept(mypiv, eps));
            if (ii == 0 && bad) *fail = kb + __ffsll((long long)bad);
        } else if (p >= 1) {
            // block row p - 1 of X: wave w forms piece j = w - 1 (wave 1 writes the diagonal block)
            const int pr = p - 1, j = wv - 1;
            if (j < pr || wv == 1) inverse_row<T>(D, rd, X, pr, min(j, pr), wv == 1);
        } else {
            pre(wv);
        }
        SDS_POTRF_MARK(0);
        __syncthreads();
        SDS_POTRF_MARK(1);
        if (*fail) break;  // uniform after the barrier
        // trailing tiles (I, J), kb + 16 <= J <= I < 64: A(I, J) -= L(I, panel) L(J, panel)^T
        const int nr = (NB - kb - 16) / 16, ntile = nr * (nr + 1) / 2;
        const int lane = tid & 63, l16 = lane & 15, l4 = lane >> 4;
        for (int q0 = wv; q0 < ntile; q0 += 4) {
            int bi = 0, rem = q0;
            while (rem > bi) rem -= ++bi;  // q0 = bi (bi + 1) / 2 + bj, bj <= bi
            const int I = kb + 16 + 16 * bi, J = kb + 16 + 16 * rem;
            acc_t acc;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = D[J + l16][I + Mfma<T>::row(l4, q)];
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                acc = Mfma<T>::mac(-D[kb + 4 * ks + l4][I + l16], D[kb + 4 * ks + l4][J + l16], acc);
#pragma unroll
            for (int q = 0; q < 4; ++q) D[J + l16][I + Mfma<T>::row(l4, q)] = acc[q];
        }
        __syncthreads();
        SDS_POTRF_MARK(2);
    }
    if (!*fail) {
        // the last factored block row P (wave w: piece w, the diagonal block alone for w = P),
        // and the identity for the block rows past it (pure padding)
        const int P = p - 1;
        if (wv <= P) inverse_row<T>(D, rd, X, P, wv, wv == P);
        for (int e = tid; e < NB * NB; e += blockDim.x) {
            const int j = e / NB, i = e - j * NB;
            if (i / 16 > P && j / 16 <= i / 16) X[j][i] = i == j ? T(1) : T(0);
        }
    }
    for (int j = nb + tid; j < NB; j += blockDim.x) rd[j] = T(1);
    __syncthreads();
    SDS_POTRF_MARK(3);
}

// after block_potrf_inv: the factored block into the panel (lower part)
// and Linv into DB (NB entries per factor column, zero past nb) for the later row solves and
// the sweeps
template <typename T>
__device__ __attribute__((always_inline)) inline void block_store(const T (*D)[NB + 1], const T (*X)[NB + 1], T* P,
                                                                  int f, int k, int nb, T* out) {
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) {
        const int j = q / NB, i = q - j * NB;
        if (j < nb && i >= j && i < nb) P[(int64_t)(k + j) * f + k + i] = D[j][i];
        // X is lower block triangular: its blocks above the diagonal blocks are never formed
        if (j < nb) out[q] = i < nb && i / 16 >= j / 16 ? X[j][i] : T(0);
    }
}

// Diagonal block of a large front's first step: its Cholesky and inverse, one workgroup per
// front. The later steps' diagonal blocks are factored inside k_big_step. task: (s, k)
template <typename T>
__global__ void __launch_bounds__(256) k_big_diag(DevSym S, const Task* __restrict__ tasks, T* __restrict__ factor,
                                                 T* __restrict__ DB, int32_t* __restrict__ status, double eps) {
    __shared__ T D[NB][NB + 1];
    __shared__ T X[NB][NB + 1];
    __shared__ T rd[NB];
    __shared__ int fail;
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a;
    if (status[s] != 0) return;
    const int nb = min(NB, w - k);
    T* P = factor + S.panel_off[s];
    // the block's sixteen loads per thread in flight together (clamped into the block), then the
    // masking
    T v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u, j = q / NB, ii = q - j * NB;
        v[u] = P[(int64_t)(k + min(j, nb - 1)) * f + k + min(ii, nb - 1)];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u, j = q / NB, ii = q - j * NB;
        D[j][ii] = (ii < nb && j < nb) ? (ii >= j ? v[u] : T(0)) : (ii == j ? T(1) : T(0));
    }
    block_potrf_inv<T>(D, rd, &fail, nb, eps, X);
    if (fail) {
        if (threadIdx.x == 0) status[s] = k + fail;
        return;
    }
    block_store<T>(D, X, P, f, k, nb, DB + (int64_t)(S.super_ptr[s] + k) * NB);
}

// Schur update of one 64x64 tile on the matrix cores: C(I, J) -= A(I, k0:k1) A(J, k0:k1)^T
// on the lower triangle of the front, A the panel columns k0..k1-1. mode 0: step update of
// the panel columns (k1 = min(k0 + NB, w)); mode 1: contribution block (k0 = 0, k1 = w).
// task: (s, I0, J0, k0). 256 threads = 4 waves, each wave a 32x32 quadrant of the tile as
// 2x2 blocks of 16x16 accumulators; K in chunks of KC staged through LDS (wave v stages rows
// v, v + 4, ... of a chunk), the next chunk's loads issued between the current chunk's
// products. EARLY: the tile's old values are read at the start (one workgroup per CU: their
// latency hides under the products) rather than at the end (several workgroups per CU, fewer
// live registers). The accumulators hold the transposed tile, so that each register covers 16
// consecutive rows of one column (coalesced loads and stores). MFMA 16x16x4 operand layout
// (checked on gfx950): lane l gives A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator
// element q of lane l is D[row(l / 16, q)][l % 16] with row = 4 (l / 16) + q for the f32
// instruction and 4 q + l / 16 for the f64 one.

template <typename T, int KC, bool EARLY>
__device__ __attribute__((always_inline)) inline void update_tile(const DevSym& S, const Task t, T* __restrict__ factor,
                                                                  T* __restrict__ cbpool,
                                                                  const int32_t* __restrict__ status, int mode,
                                                                  T (*As)[TILE], T (*Bs)[TILE], int pblk,
                                                                  T (*Dout)[NB + 1] = nullptr, int nbd = 0) {
    constexpr int PER = KC * TILE / 256;  // chunk rows per wave
    constexpr int EVERY = KC / 4 / PER;   // products per issued pair of next-chunk loads
    typedef typename Mfma<T>::acc_t acc_t;
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    if (status[s] != 0) return;
    const int I0 = t.a, J0 = t.b;
    // mode 0: K starts at column c; it spans pblk columns (a whole panel block) when the BIGK bit
    // of c is set, one step (NB) otherwise
    const int kc0 = t.c & ~BIGK;
    const int k0 = mode == 0 ? kc0 : 0;
    const int k1 = mode == 0 ? min(kc0 + ((t.c & BIGK) ? pblk : NB), w) : w;
    const int jlim = mode == 0 ? w : f;
    const int jmin = mode == 0 ? k1 : w;
    T* P = factor + S.panel_off[s];
    T* C = cbpool + (S.cb_off[s] >= 0 ? S.cb_off[s] : 0);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int l16 = lane & 15, l4 = lane >> 4;
    // element q of lane l of block (a, b) is entry (I0 + 32 wm + 16 a + l % 16, J0 + 32 wn + 16 b + row(l / 16, q))
    T old[2][2][4];
    if (EARLY)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = I0 + wm * 32 + a * 16 + l16;
                    const int j = J0 + wn * 32 + b * 16 + Mfma<T>::row(l4, q);
                    const bool in = j < jlim && j >= jmin && i < f && i >= j;
                    old[a][b][q] = in ? *front_at(P, C, f, w, i, j) : T(0);
                }
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    // wave wv stages chunk rows kk = wv + 4 u (panel column kc + kk), lanes along the tile rows
    const bool rowA = I0 + lane < f, rowB = J0 + lane < f;
    const int64_t s4 = 4 * (int64_t)f;
    const T* pA = P + (int64_t)(k0 + wv) * f + min(I0 + lane, f - 1);
    const T* pB = P + (int64_t)(k0 + wv) * f + min(J0 + lane, f - 1);
    T pa[PER], pb[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const bool kin = k0 + wv + 4 * u < k1;
        pa[u] = kin && rowA ? pA[u * s4] : T(0);
        pb[u] = kin && rowB ? pB[u * s4] : T(0);
    }
    for (int kc = k0; kc < k1; kc += KC) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            As[wv + 4 * u][lane] = pa[u];
            Bs[wv + 4 * u][lane] = pb[u];
        }
        __syncthreads();
        pA += KC * (int64_t)f;
        pB += KC * (int64_t)f;
        const int kn = min(KC, k1 - kc);
        const int kcn = kc + KC;
        if (kn == KC) {
#pragma unroll
            for (int k4 = 0; k4 < KC; k4 += 4) {
                if ((k4 / 4) % EVERY == 0) {
                    const int u = k4 / 4 / EVERY;
                    const bool kin = kcn + wv + 4 * u < k1;
                    pa[u] = kin && rowA ? pA[u * s4] : T(0);
                    pb[u] = kin && rowB ? pB[u * s4] : T(0);
                }
                const T a0 = As[k4 + l4][wm * 32 + l16], a1 = As[k4 + l4][wm * 32 + 16 + l16];
                const T b0 = Bs[k4 + l4][wn * 32 + l16], b1 = Bs[k4 + l4][wn * 32 + 16 + l16];
                acc[0][0] = Mfma<T>::mac(b0, a0, acc[0][0]);
                acc[0][1] = Mfma<T>::mac(b1, a0, acc[0][1]);
                acc[1][0] = Mfma<T>::mac(b0, a1, acc[1][0]);
                acc[1][1] = Mfma<T>::mac(b1, a1, acc[1][1]);
            }
        } else {  // the last, partial chunk (nothing left to load)
            for (int k4 = 0; k4 < kn; k4 += 4) {
                const T a0 = As[k4 + l4][wm * 32 + l16], a1 = As[k4 + l4][wm * 32 + 16 + l16];
                const T b0 = Bs[k4 + l4][wn * 32 + l16], b1 = Bs[k4 + l4][wn * 32 + 16 + l16];
                acc[0][0] = Mfma<T>::mac(b0, a0, acc[0][0]);
                acc[0][1] = Mfma<T>::mac(b1, a0, acc[0][1]);
                acc[1][0] = Mfma<T>::mac(b0, a1, acc[1][0]);
                acc[1][1] = Mfma<T>::mac(b1, a1, acc[1][1]);
            }
        }
        __syncthreads();
    }
    if (!EARLY && mode == 0) {
        // the tile's old values (panel columns only in mode 0), read in one batch from clamped
        // addresses after the products: read one at a time next to their stores, each read
        // waited for the stores before it
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = I0 + wm * 32 + a * 16 + l16;
                    const int j = J0 + wn * 32 + b * 16 + Mfma<T>::row(l4, q);
                    old[a][b][q] = P[(int64_t)min(j, w - 1) * f + min(i, f - 1)];
                }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) asm volatile("" : "+v"(old[a][b][q]));
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = I0 + wm * 32 + a * 16 + l16;
                const int j = J0 + wn * 32 + b * 16 + Mfma<T>::row(l4, q);
                const bool in = j < jlim && j >= jmin && i < f && i >= j;
                if (Dout) {
                    // a diagonal tile about to be factored: its lower nbd x nbd block goes to LDS
                    // (upper part zero, identity past nbd), rows below the block to the panel
                    const int ii = i - I0, jj = j - J0;
                    if (ii < nbd && jj < nbd) {
                        Dout[jj][ii] = in ? (EARLY || mode == 0 ? old[a][b][q] : *front_at(P, C, f, w, i, j)) - acc[a][b][q] : T(0);
                        continue;
                    }
                    if (jj >= nbd) Dout[jj][ii] = ii == jj ? T(1) : T(0);
                    else Dout[jj][ii] = T(0);
                }
                if (in) {
                    // mode 1 writes the contribution block whole (the children's entries are
                    // added to it afterwards)
                    T* d = front_at(P, C, f, w, i, j);
                    *d = (EARLY || mode == 0 ? old[a][b][q] : T(0)) - acc[a][b][q];
                }
            }
}

// MODE a template parameter: each mode gets its own register allocation (mode 0 holds the
// tile's old values; the contribution-block update, mode 1, keeps four workgroups per CU)
template <typename T, int KC, int MODE>
__global__ void __launch_bounds__(256) k_big_update_mfma(DevSym S, const Task* __restrict__ tasks,
                                                        T* __restrict__ factor, T* __restrict__ cbpool,
                                                        const int32_t* __restrict__ status, int pb) {
    __shared__ T As[KC][TILE];
    __shared__ T Bs[KC][TILE];
    update_tile<T, KC, false>(S, tasks[blockIdx.x], factor, cbpool, status, MODE, As, Bs, pb);
}

// the rows [k + 64, k + 128) of a step's columns, kept in LN by k_big_step, into the panel.
// task: (s, k)
template <typename T>
__device__ inline void copy_ln(const DevSym& S, const Task t, T* __restrict__ factor, const T* __restrict__ LN,
                               const int32_t* __restrict__ status) {
    const int s = t.s, f = S.fnrows[s], k = t.a, J = k + NB;
    if (status[s] != 0) return;
    T* P = factor + S.panel_off[s];
    const T* src = LN + (int64_t)(S.super_ptr[s] + k) * NB;
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int c = e / NB, r = e - c * NB;
        if (J + r < f) P[(int64_t)(k + c) * f + J + r] = src[e];
    }
}

// One step b (columns [k, k + 64)) of a large front for a tile of 64 rows [I0, I0 + 64) below
// the step's diagonal block, on the matrix cores, one workgroup of 4 waves (16 rows each):
//  1. row solve: L(I, b) = A(I, b) Linv_b^T (Linv_b in DB, staged in LDS; the A fragments are
//     all loaded before the MFMAs, so the rows are replaced in place);
//  2. look-ahead, when a next block b + 1 exists (J = k + 64): the tile's rows of the next
//     block's columns get C(I, J) -= L(I, b) L(J, b)^T, and also -= L(I, b-1) L(J, b-1)^T
//     when b > 0 and I0 > J (the steps before b-1 reach block b + 1 through the trailing-update
//     tiles, which cover the blocks from three past their step). L(J, b) is recomputed by
//     every workgroup from A(J, b);
//  3. the workgroup whose tile is the next diagonal block (I0 = J) factors it (Cholesky and
//     inverse, as k_big_diag) for step b + 1. Its tile got the step b-1 part one launch
//     earlier (4.), so its chain holds the step's own product only;
//  4. the workgroup of the rows after (I0 = J + 64) also updates its own diagonal tile, block
//     b + 2, with steps b and b-1; the trailing-update tiles of step b-1 leave that tile out.
// The rows J of the step's columns are read (unsolved) by every workgroup, so the workgroup
// that solves them keeps them in LN (NB entries per factor column) instead of the panel.
// The next step's launch also carries this step's copies of those rows (copy_ln) and its
// trailing update of the blocks three on (update_tile, mode 0): they touch no block that
// step reads or writes, and they fill the machine while the step's chain runs.
// Panel blocks (pb > 0, a multiple of NB): the chain above runs inside blocks of pb columns
// [K0, K1) only. The look-ahead stops at K1; the columns from K1 on get the block's whole
// update with K = pb from k_big_block (the next block's first strip, whose launch also factors
// the next diagonal block) and from update tiles with the BIGK flag that ride in the next
// block's step launches (the strips after it). A block's first step has no previous step in
// the chain: k_big_block applied everything before K0.
// tasks: nstep steps (s, k, I0), then ncopy copies (s, k), then update tiles (s, I0, J0, k)
// OCC: workgroups per CU the kernel is built for (1: registers unbounded; 2: at most 256 VGPRs
// per lane, for the launches with many workgroups). The two LDS tiles (67 KiB in FP64) allow two.
template <typename T, int OCC>
__global__ void __launch_bounds__(256, OCC) k_big_step(DevSym S, const Task* __restrict__ tasks,
                                                      T* __restrict__ factor, T* __restrict__ cbpool,
                                                      T* __restrict__ DB, T* __restrict__ LN,
                                                      int32_t* __restrict__ status, double eps, int nstep,
                                                      int ncopy, int pb) {
    typedef typename Mfma<T>::acc_t acc_t;
    // Bs[kk][c] = Linv_b(c, kk); later L(J + r, b - 1) as Bs[r][kk]; later L(I0 + r, k + kk) (the
    // rows after next) or D
    __shared__ T Bs[NB][NB + 1];
    __shared__ T Ln[NB][NB + 1];  // Ln[r][kk] = L(J + r, k + kk); later L(I0 + r, k - NB + kk) or Linv_{b+1}
    __shared__ T rd[NB];
    __shared__ int fail;
    const Task t = tasks[blockIdx.x];
    if ((int)blockIdx.x >= nstep + ncopy) {
        update_tile<T, NB, true>(S, t, factor, cbpool, status, 0, reinterpret_cast<T(*)[TILE]>(&Bs[0][0]),
                                 reinterpret_cast<T(*)[TILE]>(&Ln[0][0]), pb);
        return;
    }
    if ((int)blockIdx.x >= nstep) {
        copy_ln<T>(S, t, factor, LN, status);
        return;
    }
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a, I0 = t.b;
    if (status[s] != 0) return;
    SDS_STEP_MARK(0);
    const int nb = min(NB, w - k);
    // the panel block of this step: [K0, K1)
    const int K0 = pb > 0 ? k - k % pb : 0, K1 = pb > 0 ? min(w, K0 + pb) : w;
    const bool la = k + NB < K1;  // a next block exists in the chain (then nb = NB)
    const int J = k + NB, nbn = la ? min(NB, K1 - J) : 0;
    T* P = factor + S.panel_off[s];
    const int cg = S.super_ptr[s];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
    const T* Li = DB + (int64_t)(cg + k) * NB;
    const int rI = I0 + 16 * wv + l16, rJ = J + 16 * wv + l16;
    // the diagonal workgroup's rows are the rows J: it reads them once
    const bool two = la && I0 != J;
    // every load of the phase is issued (from clamped, in-bounds addresses) before any of them is
    // used: a use next to its load makes the compiler wait for each load before the next one.
    // Linv_b (its first nb columns), the step's columns of the tile's rows (and of the rows J)
    T lv[16], aI[16], aJ[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) lv[u] = Li[min(tid + 256 * u, nb * NB - 1)];
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) aI[ks] = P[(int64_t)(k + min(4 * ks + l4, nb - 1)) * f + min(rI, f - 1)];
    if (two)
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) aJ[ks] = P[(int64_t)(k + min(4 * ks + l4, nb - 1)) * f + min(rJ, f - 1)];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int e = tid + 256 * u, kk = e / NB, c = e - kk * NB;
        Bs[kk][c] = kk < nb && c < nb ? lv[u] : T(0);
    }
#pragma unroll
    for (int ks = 0; ks < 16; ++ks) aI[ks] = rI < f && 4 * ks + l4 < nb ? aI[ks] : T(0);
    if (two)
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) aJ[ks] = rJ < f && 4 * ks + l4 < nb ? aJ[ks] : T(0);
    __syncthreads();
    SDS_STEP_MARK(1);
    // the look-ahead's operands that do not depend on this step, read after the barrier (which
    // waits only for Linv_b and the step's rows) so that their latency hides under the row
    // solve: the previous step's rows J and I0, and the tile's old values of the next block's
    // columns (all final before this launch: the trailing-update tiles of this launch touch
    // neither)
    const bool prev = k >= K0 + NB;
    const int kp = max(k - NB, 0);
    T pj[16], ap[16];
    acc_t accC[4];  // transposed tile: element q of lane l is C(I0 + 16 wv + l % 16, J + 16 nt + row(l / 16, q))
    if (two)  // then kp + 63 < w: the columns read lie in the panel
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = tid + 256 * u, kk = e / NB, r = e - kk * NB;
            pj[u] = P[(int64_t)(kp + kk) * f + min(J + r, f - 1)];
        }
    if (two)
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) ap[ks] = P[(int64_t)(kp + 4 * ks + l4) * f + min(rI, f - 1)];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = J + 16 * nt + Mfma<T>::row(l4, q);
            accC[nt][q] = P[(int64_t)min(j, w - 1) * f + min(rI, f - 1)];
        }
    // (masked after the row solve, so that the loads stay in flight under it)
    __builtin_amdgcn_sched_barrier(0);
    // the row solve as L(I0, b)^T = Linv_b A(I0, b)^T: transposed products, so that each
    // accumulator register holds 16 consecutive rows of one column (coalesced stores); element
    // q of lane l is L(I0 + 16 wv + l % 16, k + 16 nt + row(l / 16, q))
    acc_t accI[4], accJ[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) accI[nt] = accJ[nt] = acc_t{};
    if (two) {
#pragma unroll
        for (int ks = 0; ks < 16; ++ks)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const T a = Bs[4 * ks + l4][16 * nt + l16];
                accI[nt] = Mfma<T>::mac(a, aI[ks], accI[nt]);
                accJ[nt] = Mfma<T>::mac(a, aJ[ks], accJ[nt]);
            }
    } else {
        // the diagonal workgroup's rows are the rows J (or there is no next block): one product
#pragma unroll
        for (int ks = 0; ks < 16; ++ks)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) accI[nt] = Mfma<T>::mac(Bs[4 * ks + l4][16 * nt + l16], aI[ks], accI[nt]);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) accJ[nt] = accI[nt];
    }
    const int r = 16 * wv + l16;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = 16 * nt + Mfma<T>::row(l4, q);
            if (I0 + r < f && c < nb) {
                if (la && I0 == J)
                    LN[(int64_t)(cg + k + c) * NB + r] = accI[nt][q];
                else
                    P[(int64_t)(k + c) * f + I0 + r] = accI[nt][q];
            }
            if (la) Ln[r][c] = J + r < f ? accJ[nt][q] : T(0);
        }
    __builtin_amdgcn_sched_barrier(0);
    if (two)
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = tid + 256 * u, rr = e - (e / NB) * NB;
            pj[u] = prev && J + rr < f ? pj[u] : T(0);
        }
    if (two)
#pragma unroll
        for (int ks = 0; ks < 16; ++ks) ap[ks] = prev && rI < f ? ap[ks] : T(0);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = J + 16 * nt + Mfma<T>::row(l4, q);
            accC[nt][q] = rI < f && j < J + nbn && rI >= j ? accC[nt][q] : T(0);
        }
    SDS_STEP_MARK(2);
    if (!la) return;
    // the previous step's rows of the next block into Bs, for its part of the look-ahead (the
    // diagonal tile has that part already: the step before applied it, below)
    __syncthreads();
    const bool pv = prev && two;
    if (pv) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int e = tid + 256 * u, kk = e / NB, rr = e - kk * NB;
            Bs[rr][kk] = pj[u];
        }
        __syncthreads();
    }
    SDS_STEP_MARK(3);
    if (I0 == J && nbn == NB) {
        // the next diagonal block, a full one: its old values (lower part) go to D, and its
        // update by step b, D -= L(J, b) L(J, b)^T, is made per 16-column block from Ln (rows J
        // of step b, all in LDS): the first column block by the four waves before the
        // factorization, the others by waves 1-3 while wave 0 factors the first 16 columns. Each
        // block's products run in the order of the whole tile's, so the values are the same.
        T(*D)[NB + 1] = Bs;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = 16 * nt + Mfma<T>::row(l4, q);
                D[j][r] = r >= j ? accC[nt][q] : T(0);
            }
        __syncthreads();
        // block (rows 16 rb.., columns 16 cb..) of the tile, rb >= cb
        auto upd = [&](int rb, int cb) {
            acc_t acc;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = D[16 * cb + Mfma<T>::row(l4, q)][16 * rb + l16];
#pragma unroll
            for (int ntp = 0; ntp < 4; ++ntp)
#pragma unroll
                for (int qp = 0; qp < 4; ++qp) {
                    const int kx = 16 * ntp + Mfma<T>::row(l4, qp);
                    acc = Mfma<T>::mac(-Ln[16 * cb + l16][kx], Ln[16 * rb + l16][kx], acc);
                }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = 16 * rb + l16, j = 16 * cb + Mfma<T>::row(l4, q);
                D[j][i] = i >= j ? acc[q] : T(0);
            }
        };
        upd(wv, 0);
        __syncthreads();
        SDS_STEP_MARK(5);
        // waves 1-3: the six blocks of column blocks 1-3 on or below the diagonal, two each
        // (wave 1: (1, 1) and (3, 3); wave 2: (2, 1) and (2, 2); wave 3: (3, 1) and (3, 2))
        block_potrf_inv<T>(D, rd, &fail, NB, eps, Ln, [&](int w) {
            upd(w, 1);
            if (w == 1)
                upd(3, 3);
            else if (w == 2)
                upd(2, 2);
            else
                upd(3, 2);
        });
        SDS_STEP_MARK(6);
        if (fail) {
            if (tid == 0) status[s] = J + fail;
            return;
        }
        SDS_STEP_MARK(7);
        block_store<T>(D, Ln, P, f, J, NB, DB + (int64_t)(cg + J) * NB);
        SDS_STEP_MARK(8);
        return;
    }
    // C(I0.., J..)^T -= L(J, b) L(I0, b)^T (+ L(J, b - 1) L(I0, b - 1)^T off the diagonal tile);
    // the K index of the first product runs over accI's registers (element q of accI[nt'] is
    // L(I0 + r, k + 16 nt' + row(l / 16, q)), the B operand of this wave's rows as it stands)
#pragma unroll
    for (int ntp = 0; ntp < 4; ++ntp)
#pragma unroll
        for (int qp = 0; qp < 4; ++qp) {
            const int kx = 16 * ntp + Mfma<T>::row(l4, qp);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) accC[nt] = Mfma<T>::mac(-Ln[16 * nt + l16][kx], accI[ntp][qp], accC[nt]);
        }
    if (pv)
#pragma unroll
        for (int ks = 0; ks < 16; ++ks)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) accC[nt] = Mfma<T>::mac(-Bs[16 * nt + l16][4 * ks + l4], ap[ks], accC[nt]);
    SDS_STEP_MARK(4);
    if (I0 != J) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = rI, j = J + 16 * nt + Mfma<T>::row(l4, q);
                if (i < f && j < J + nbn) P[(int64_t)j * f + i] = accC[nt][q];
            }
        const int I = J + NB;  // the block after next
        if (I0 != I || I >= K1) return;
        // this workgroup's rows are the block after next: its diagonal tile gets steps b and
        // b - 1 here (the trailing-update tiles of step b - 1 in this launch leave it out), so
        // that the next step's diagonal workgroup only applies its own step
        const int nb3 = min(NB, K1 - I);
        acc_t accD[4];  // element q of lane l is C(I + 16 wv + l % 16, I + 16 nt + row(l / 16, q))
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                accD[nt][q] = P[(int64_t)min(I + 16 * nt + Mfma<T>::row(l4, q), w - 1) * f + min(rI, f - 1)];
        __builtin_amdgcn_sched_barrier(0);  // the sixteen loads in flight together
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = rI, j = I + 16 * nt + Mfma<T>::row(l4, q);
                accD[nt][q] = i < f && j < I + nb3 && i >= j ? accD[nt][q] : T(0);
            }
        __syncthreads();  // Bs and Ln are free: L(I, b) and L(I, b - 1) into them as rows
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[r][16 * nt + Mfma<T>::row(l4, q)] = accI[nt][q];
        if (prev)
#pragma unroll
            for (int ks = 0; ks < 16; ++ks) Ln[r][4 * ks + l4] = ap[ks];
        __syncthreads();
        // step b: the K index runs over accI's registers, the rows of L(I, b) from Bs
#pragma unroll
        for (int ntp = 0; ntp < 4; ++ntp)
#pragma unroll
            for (int qp = 0; qp < 4; ++qp) {
                const int kx = 16 * ntp + Mfma<T>::row(l4, qp);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) accD[nt] = Mfma<T>::mac(-Bs[16 * nt + l16][kx], accI[ntp][qp], accD[nt]);
            }
        if (prev)
#pragma unroll
            for (int ks = 0; ks < 16; ++ks)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) accD[nt] = Mfma<T>::mac(-Ln[16 * nt + l16][4 * ks + l4], ap[ks], accD[nt]);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = rI, j = I + 16 * nt + Mfma<T>::row(l4, q);
                if (i < f && j < I + nb3 && i >= j) P[(int64_t)j * f + i] = accD[nt][q];
            }
        return;
    }
    // the next diagonal block: factor and invert it for step b + 1 (a narrower last block
    // leaves rows below it in this tile: those are stored as updated)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = r, j = 16 * nt + Mfma<T>::row(l4, q);
            if (i >= nbn && J + i < f && j < nbn) P[(int64_t)(J + j) * f + J + i] = accC[nt][q];
        }
    __syncthreads();
    T(*D)[NB + 1] = Bs;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = r, j = 16 * nt + Mfma<T>::row(l4, q);
            D[j][i] = (i < nbn && j < nbn) ? (i >= j ? accC[nt][q] : T(0)) : (i == j ? T(1) : T(0));
        }
    SDS_STEP_MARK(5);
    block_potrf_inv<T>(D, rd, &fail, nbn, eps, Ln);
    SDS_STEP_MARK(6);
    if (fail) {
        if (tid == 0) status[s] = J + fail;
        return;
    }
    SDS_STEP_MARK(7);
    block_store<T>(D, Ln, P, f, J, nbn, DB + (int64_t)(cg + J) * NB);
    SDS_STEP_MARK(8);
}


// The first strip of the next panel block: tiles (s, I0, J0, K0 | BIGK) with J0 in [K1, K1 + pb)
// get C(I, J) -= L(I, K0:K1) L(J, K0:K1)^T with K = pb (K1 = K0 + pb), on the matrix cores
// (update_tile). The first ndiag tasks are the next block's diagonal tiles (I0 = J0 = K1): their
// workgroup keeps the updated block in LDS and factors and inverts it (as k_big_diag) for the
// next block's step chain; its rows below a narrower last block go to the panel.
template <typename T>
__global__ void __launch_bounds__(256, 2) k_big_block(DevSym S, const Task* __restrict__ tasks,
                                                     T* __restrict__ factor, T* __restrict__ cbpool,
                                                     T* __restrict__ DB, int32_t* __restrict__ status, double eps,
                                                     int pb, int ndiag) {
    __shared__ T D[NB][NB + 1];
    __shared__ T X[NB][NB + 1];
    __shared__ T rd[NB];
    __shared__ int fail;
    const Task t = tasks[blockIdx.x];
    T(*As)[TILE] = reinterpret_cast<T(*)[TILE]>(&D[0][0]);
    T(*Bs)[TILE] = reinterpret_cast<T(*)[TILE]>(&X[0][0]);
    if ((int)blockIdx.x >= ndiag) {
        update_tile<T, NB, true>(S, t, factor, cbpool, status, 0, As, Bs, pb);
        return;
    }
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.b;
    if (status[s] != 0) return;
    const int nb = min(NB, w - k);
    update_tile<T, NB, true>(S, t, factor, cbpool, status, 0, As, Bs, pb, D, nb);
    __syncthreads();
    block_potrf_inv<T>(D, rd, &fail, nb, eps, X);
    if (fail) {
        if (threadIdx.x == 0) status[s] = k + fail;
        return;
    }
    block_store<T>(D, X, factor + S.panel_off[s], f, k, nb, DB + (int64_t)(S.super_ptr[s] + k) * NB);
}

// per-front statistics of a Cholesky (src/numeric/factorize.jl:13-48) and, per block, the
// smallest failed factor column (1-based, INT64_MAX = none) into part[blockIdx.x]
__global__ void __launch_bounds__(256) k_chol_stats(DevSym S, const int32_t* __restrict__ status,
                                                   int64_t* __restrict__ stats, int64_t* __restrict__ part) {
    __shared__ int64_t best[256];
    int64_t m = INT64_MAX;
    for (int64_t s = blockIdx.x * 256 + threadIdx.x; s < S.ns; s += (int64_t)gridDim.x * 256) {
        const int64_t fi = status[s];
        int64_t* st = stats + s * NF;
        st[0] = fi == 0 ? (int64_t)S.fncols[s] : fi - 1;
        st[1] = st[2] = st[3] = st[4] = 0;
        st[5] = fi;
        if (fi > 0) m = min(m, (int64_t)S.super_ptr[s] + fi);
    }
    best[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) best[threadIdx.x] = min(best[threadIdx.x], best[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = best[0];
}

// info[0] = the smallest of the n block minima, 0 for none (one block)
__global__ void __launch_bounds__(256) k_info_final(const int64_t* __restrict__ part, int n,
                                                   int64_t* __restrict__ info) {
    __shared__ int64_t best[256];
    int64_t m = INT64_MAX;
    for (int b = threadIdx.x; b < n; b += 256) m = min(m, part[b]);
    best[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) best[threadIdx.x] = min(best[threadIdx.x], best[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) info[0] = best[0] == INT64_MAX ? 0 : best[0];
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Schur complement mode: the root front s (the Schur block) is assembled, not factored

// its factor columns: the identity local pivot order (global columns, as the pivoting kernels
// leave piv), unit 1x1 pivots with d = 1, zero statistics (src/numeric/schur.jl:15-28)
template <typename T>
__global__ void k_schur_meta(DevSym S, int s, int32_t* __restrict__ piv, int8_t* __restrict__ kind,
                             T* __restrict__ dvec, int64_t* __restrict__ stats) {
    const int c0 = S.super_ptr[s], w = S.fncols[s];
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < max(w, NF); q += gridDim.x * blockDim.x) {
        if (q < w) {
            if (piv) piv[c0 + q] = c0 + q;
            if (kind) kind[c0 + q] = KIND_1X1;
            if (dvec) {
                dvec[c0 + q] = T(1);
                dvec[S.n + c0 + q] = T(0);
            }
        }
        if (q < NF) stats[(int64_t)s * NF + q] = 0;
    }
}

// the Schur complement S (ns x ns, column-major, leading dimension lds, Float64) from the root's
// panel: the symmetric structures assemble the lower triangle (the upper is mirrored); "G" (full:
// a large LU front assembles its whole fully summed square in the L panel, U^T holds only the rows
// past it, and the root has none) the whole square
template <typename T>
__global__ void k_schur_export(DevSym S, int s, const T* __restrict__ factor, int full, double* __restrict__ out,
                               int64_t lds) {
    const int ns = S.fncols[s];
    const T* P = factor + S.panel_off[s];
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < (int64_t)ns * ns;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(q / ns), i = (int)(q - (int64_t)j * ns);
        out[(int64_t)j * lds + i] = (double)(i >= j || full ? P[(int64_t)j * ns + i] : P[(int64_t)i * ns + j]);
    }
}

// ---------------------------------------------------------------------------
// launchers

template <typename T>
void launch_front_small_chol(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                             const double* nz, T* factor, T* cbpool, T* DB, int32_t* status, double eps) {
    size_t lds = (size_t)fmax * fmax * sizeof(T);
    hipLaunchKernelGGL(k_front_small_chol<T>, dim3(nfronts), dim3(256), lds, st, S, fronts, nz, factor, cbpool, DB,
                       status, eps);
}

template <typename T>
void launch_subtree_chol(hipStream_t st, const DevSym& S, const int32_t* sub_ptr, const int32_t* sub_nodes, int first,
                         int count, int lds, const int32_t* lfront, const int32_t* lcb, const double* nz, T* factor,
                         T* cbpool, T* DB, int32_t* status, double eps) {
    hipLaunchKernelGGL(k_subtree_chol<T>, dim3(count), dim3(256), (size_t)lds, st, S, sub_ptr, sub_nodes, first,
                       lfront, lcb, nz, factor, cbpool, DB, status, eps);
}

template <typename T>
void launch_front_col_chol(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                           const double* nz, T* factor, T* cbpool, T* DB, int32_t* status, double eps) {
    size_t lds = (size_t)fmax * fmax * sizeof(T);
    hipLaunchKernelGGL(k_front_col_chol<T>, dim3(nfronts), dim3(256), lds, st, S, fronts, nz, factor, cbpool, DB,
                       status, eps);
}

template <typename T>
void launch_front_tiny_chol(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, const double* nz,
                            T* factor, T* cbpool, T* DB, int32_t* status, double eps) {
    hipLaunchKernelGGL(k_front_tiny_chol<T>, dim3((nfronts + 3) / 4), dim3(256), 0, st, S, fronts, nfronts, nz, factor,
                       cbpool, DB, status, eps);
}

template <typename T>
void launch_big_zero_scatter(hipStream_t st, const DevSym& S, const Task* tasks, int n, const double* nz, T* factor,
                             T* cbpool) {
    hipLaunchKernelGGL(k_big_zero_scatter<T>, dim3(n), dim3(256), 0, st, S, tasks, nz, factor, cbpool);
}

template <typename T>
void launch_big_extend_add(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* cbpool) {
    hipLaunchKernelGGL(k_big_extend_add<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool);
}

template <typename T>
void launch_big_diag(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* DB, int32_t* status,
                     double eps) {
    hipLaunchKernelGGL(k_big_diag<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, DB, status, eps);
}

template <typename T>
void launch_big_step(hipStream_t st, const DevSym& S, const Task* tasks, int n, int nstep, int ncopy, T* factor,
                     T* cbpool, T* DB, T* LN, int32_t* status, double eps, int occ, int pb) {
    if (occ == 2)
        hipLaunchKernelGGL((k_big_step<T, 2>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, DB, LN, status, eps,
                           nstep, ncopy, pb);
    else
        hipLaunchKernelGGL((k_big_step<T, 1>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, DB, LN, status, eps,
                           nstep, ncopy, pb);
}

template <typename T>
void launch_big_block(hipStream_t st, const DevSym& S, const Task* tasks, int n, int ndiag, T* factor, T* cbpool,
                      T* DB, int32_t* status, double eps, int pb) {
    hipLaunchKernelGGL(k_big_block<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, DB, status, eps, pb,
                       ndiag);
}

template <typename T>
void launch_big_update(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* cbpool,
                       const int32_t* status, int mode, int kc, int pb) {
    if (mode == 0) {
        if (kc == 64)
            hipLaunchKernelGGL((k_big_update_mfma<T, 64, 0>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, status, pb);
        else if (kc == 16)
            hipLaunchKernelGGL((k_big_update_mfma<T, 16, 0>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, status, pb);
        else
            hipLaunchKernelGGL((k_big_update_mfma<T, 32, 0>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, status, pb);
    } else {
        if (kc == 64)
            hipLaunchKernelGGL((k_big_update_mfma<T, 64, 1>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, status, pb);
        else if (kc == 16)
            hipLaunchKernelGGL((k_big_update_mfma<T, 16, 1>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, status, pb);
        else
            hipLaunchKernelGGL((k_big_update_mfma<T, 32, 1>), dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, status, pb);
    }
}

template <typename T>
void launch_schur_meta(hipStream_t st, const DevSym& S, int s, int w, int32_t* piv, int8_t* kind, T* dvec,
                       int64_t* stats) {
    hipLaunchKernelGGL(k_schur_meta<T>, dim3((std::max(w, NF) + 255) / 256), dim3(256), 0, st, S, s, piv, kind, dvec,
                       stats);
}

template <typename T>
void launch_schur_export(hipStream_t st, const DevSym& S, int s, int ns, const T* factor, int full, double* out,
                         int64_t lds) {
    const int64_t total = (int64_t)ns * ns;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 4096));
    hipLaunchKernelGGL(k_schur_export<T>, dim3(blocks), dim3(256), 0, st, S, s, factor, full, out, lds);
}

void launch_chol_stats(hipStream_t st, const DevSym& S, const int32_t* status, int64_t* stats, int64_t* info) {
    // info[8 ..] holds the block minima (INFO_BLOCKS entries)
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(INFO_BLOCKS, (S.ns + 255) / 256));
    hipLaunchKernelGGL(k_chol_stats, dim3(blocks), dim3(256), 0, st, S, status, stats, info + 8);
    hipLaunchKernelGGL(k_info_final, dim3(1), dim3(256), 0, st, info + 8, blocks, info);
}

void launch_info_final(hipStream_t st, int blocks, int64_t* info) {
    hipLaunchKernelGGL(k_info_final, dim3(1), dim3(256), 0, st, info + 8, blocks, info);
}

void launch_fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v) {
    if (n <= 0) return;
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, st, p, n, v);
}

#define INST(T)                                                                                                   \
    template void launch_front_small_chol<T>(hipStream_t, const DevSym&, const int32_t*, int, int, const double*, \
                                             T*, T*, T*, int32_t*, double);                                       \
    template void launch_subtree_chol<T>(hipStream_t, const DevSym&, const int32_t*, const int32_t*, int, int, int, \
                                         const int32_t*, const int32_t*, const double*, T*, T*, T*, int32_t*,     \
                                         double);                                                                  \
    template void launch_schur_meta<T>(hipStream_t, const DevSym&, int, int, int32_t*, int8_t*, T*, int64_t*);    \
    template void launch_schur_export<T>(hipStream_t, const DevSym&, int, int, const T*, int, double*, int64_t); \
    template void launch_front_col_chol<T>(hipStream_t, const DevSym&, const int32_t*, int, int, const double*, T*, \
                                           T*, T*, int32_t*, double);                                             \
    template void launch_front_tiny_chol<T>(hipStream_t, const DevSym&, const int32_t*, int, const double*, T*, T*, \
                                            T*, int32_t*, double);                                                \
    template void launch_big_zero_scatter<T>(hipStream_t, const DevSym&, const Task*, int, const double*, T*, T*); \
    template void launch_big_extend_add<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*);                 \
    template void launch_big_diag<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, int32_t*, double);     \
    template void launch_big_step<T>(hipStream_t, const DevSym&, const Task*, int, int, int, T*, T*, T*, T*,      \
                                     int32_t*, double, int, int);                                                  \
    template void launch_big_block<T>(hipStream_t, const DevSym&, const Task*, int, int, T*, T*, T*, int32_t*,    \
                                      double, int);                                                                \
    template void launch_big_update<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, const int32_t*, int, \
                                       int, int);
INST(double)
INST(float)
#undef INST

}  // namespace sds
