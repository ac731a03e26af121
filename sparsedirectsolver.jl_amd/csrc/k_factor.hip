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
#include "k_front.h"
#include "sds_common.h"
#include "sds_kernels.h"

namespace sds {

// Cholesky pivot acceptance on the FP64 value of the pivot (src/numeric/front.jl:46-50):
// a pivot in [0, eps] is raised to eps (cuDSS pivot_epsilon), negative or NaN fails.
__device__ inline bool chol_accept(double a, double eps) { return a > eps || (eps > 0.0 && a >= 0.0); }

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
                                                         T* __restrict__ cbpool, int32_t* __restrict__ status,
                                                         double eps) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    __shared__ int32_t st;
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) st = 0;
    assemble_dense<T>(S, s, nz, cbpool, F, f, 0);
    // right-looking Cholesky of the w pivot columns over the whole front
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
        const T d = F[j * f + j];
        for (int i = j + 1 + tid; i < f; i += nt) F[j * f + i] /= d;
        __syncthreads();
        const int r = f - j - 1;
        for (int q = tid; q < r * r; q += nt) {
            const int kk = q / r, ii = q - kk * r;
            if (ii < kk) continue;
            F[(j + 1 + kk) * f + j + 1 + ii] -= F[j * f + j + 1 + ii] * F[j * f + j + 1 + kk];
        }
        __syncthreads();
    }
    T* P = factor + S.panel_off[s];
    for (int q = tid; q < f * w; q += nt) {
        const int j = q / f, i = q - j * f;
        P[q] = i >= j ? F[q] : T(0);
    }
    const int64_t co = S.cb_off[s];
    if (co >= 0) {
        T* C = cbpool + co;
        for (int q = tid; q < m * m; q += nt) {
            const int jj = q / m, ii = q - jj * m;
            if (ii >= jj) C[q] = F[(w + jj) * f + w + ii];
        }
    }
    if (tid == 0) status[s] = st;
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
        } else {
            for (int i = tid; i < m; i += nt) C[(int64_t)(j - w) * m + i] = T(0);
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

// extend-add of one child per front (the launch covers child rank r of every front);
// task: (s, child, jj0, njj)
template <typename T>
__global__ void __launch_bounds__(256) k_big_extend_add(DevSym S, const Task* __restrict__ tasks,
                                                       T* __restrict__ factor, T* __restrict__ cbpool) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    const int c = t.a, mc = S.fnrows[c] - S.fncols[c];
    T* P = factor + S.panel_off[s];
    T* Cs = cbpool + (S.cb_off[s] >= 0 ? S.cb_off[s] : 0);
    const T* Cc = cbpool + S.cb_off[c];
    const int32_t* ri = S.relind + S.relind_ptr[c];
    const int jend = min(t.b + t.c, mc);
    constexpr int U = 8;  // entries per thread in flight: all loads issued before the stores
    for (int jj = t.b; jj < jend; ++jj) {
        const int pj = ri[jj];
        for (int ii0 = jj + threadIdx.x; ii0 < mc; ii0 += 256 * U) {
            T v[U], old[U];
            T* dst[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int ii = ii0 + 256 * u;
                dst[u] = ii < mc ? front_at(P, Cs, f, w, ri[ii], pj) : nullptr;
                v[u] = ii < mc ? Cc[(int64_t)jj * mc + ii] : T(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) old[u] = dst[u] ? *dst[u] : T(0);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (dst[u]) *dst[u] = old[u] + v[u];
        }
    }
}

// Block step of a large front: Cholesky of the (updated) diagonal block [k, k + nb) and the
// triangular solve of a tile of PANEL_ROWS panel rows below it, X <- X L_bb^-T. Every
// workgroup of the step factors the 64 x 64 block in LDS itself (two barriers per column:
// each lane recomputes the pivot's square root, no broadcast needed), so the step is one
// launch; the task with c = 1 keeps the factored block in DB (NB entries per factor column)
// and records a failed pivot. The panel's own diagonal block keeps its unfactored values
// until the factorization's last launch copies DB into it: the blocks are read by nobody in
// between, and no workgroup of the step writes what another one reads.
// Each thread then solves its row in registers, right-looking (column jj scaled by the
// pivot's reciprocal, then subtracted from the later columns: 63 independent updates); the
// row is loaded and stored straight from the panel (consecutive threads, consecutive rows).
// task: (s, k, row0, c)
template <typename T>
__global__ void __launch_bounds__(PANEL_ROWS) k_big_panel(DevSym S, const Task* __restrict__ tasks,
                                                         T* __restrict__ factor, T* __restrict__ DB,
                                                         int32_t* __restrict__ status, double eps) {
    __shared__ T D[NB][NB + 1];  // D[j][i] = L(k + i, k + j)
    __shared__ T rd[NB];         // reciprocals of the pivots
    __shared__ int fail;
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], k = t.a, row0 = t.b;
    if (status[s] != 0) return;
    const int nb = min(NB, w - k);
    T* P = factor + S.panel_off[s];
    const int tid = threadIdx.x;
    const int i = row0 + tid;
    // the diagonal block (lower part, identity past nb)
#pragma unroll
    for (int q = tid; q < NB * NB; q += PANEL_ROWS) {
        const int j = q / NB, ii = q - j * NB;
        D[j][ii] = (ii < nb && j < nb) ? (ii >= j ? P[(int64_t)(k + j) * f + k + ii] : T(0)) : (ii == j ? T(1) : T(0));
    }
    if (tid == 0) fail = 0;
    __syncthreads();
    // blocked Cholesky of D: per 16-column panel, wave 0 factors the panel's rows in registers
    // (lane i holds row kb + i of the panel; the pivot and the column entries move between
    // lanes by v_readlane, no barrier inside the panel; 1/sqrt is the hardware estimate with
    // two Newton steps), then all four waves apply the panel to the trailing block from LDS.
    const int ii = tid & (NB - 1), part = tid / NB;
    static_assert(PANEL_ROWS == 4 * NB, "four waves of NB rows");
    for (int kb = 0; kb < nb; kb += 16) {
        if (part == 0) {
            const int row = kb + ii;
            T p[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) p[c] = row < NB ? D[kb + c][row] : T(0);
            int bad = 0;
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const double piv = (double)lane_bcast(p[c], c);
                if (kb + c < nb && !chol_accept(piv, eps) && !bad) bad = kb + c + 1;
                const T r = T(rsqrt_nr(fmax(piv, eps)));
                const T d = T(fmax(piv, eps)) * r;
                p[c] = ii > c ? p[c] * r : (ii == c ? d : p[c]);
                if (ii == 0) rd[kb + c] = r;
#pragma unroll
                for (int cc = c + 1; cc < 16; ++cc) p[cc] = fma(-p[c], lane_bcast(p[c], cc), p[cc]);
            }
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (row < NB) D[kb + c][row] = p[c];
            if (ii == 0 && bad) fail = bad;
        }
        __syncthreads();
        if (fail) break;  // uniform after the barrier
        const int j0 = kb + 16;
        if (ii >= j0) {
            T lr[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) lr[c] = D[kb + c][ii];
            for (int jj = j0 + part; jj <= ii; jj += 4) {
                T acc = D[jj][ii];
#pragma unroll
                for (int c = 0; c < 16; ++c) acc = fma(-lr[c], D[kb + c][jj], acc);
                D[jj][ii] = acc;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (fail) {
        if (t.c && tid == 0) status[s] = k + fail;
        return;
    }
    for (int j = nb + tid; j < NB; j += PANEL_ROWS) rd[j] = T(1);
    if (t.c) {
        T* db = DB + (int64_t)(S.super_ptr[s] + k) * NB;
        for (int q = tid; q < nb * NB; q += PANEL_ROWS) db[q] = D[q / NB][q % NB];
    }
    __syncthreads();
    if (i >= f) return;
    // this thread's row. Per-column guards on the register row make the compiler spill it,
    // so columns past nb load column nb - 1 (they solve to garbage that nothing reads: the
    // padding of D keeps them out of the real columns) and the stores run from the last
    // column down, so that column nb - 1 is written last with its real value.
    T x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = P[(int64_t)(k + min(c, nb - 1)) * f + i];
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
        x[jj] = x[jj] * rd[jj];
        const T xj = x[jj];
#pragma unroll
        for (int kk = jj + 1; kk < NB; ++kk) x[kk] = fma(-xj, D[jj][kk], x[kk]);
    }
#pragma unroll
    for (int c = NB - 1; c >= 0; --c) P[(int64_t)(k + min(c, nb - 1)) * f + i] = x[c];
}

// Schur update of one 64x64 tile on the matrix cores: C(I, J) -= A(I, k0:k1) A(J, k0:k1)^T
// on the lower triangle of the front, A the panel columns k0..k1-1. mode 0: step update of
// the panel columns (k1 = min(k0 + NB, w)); mode 1: contribution block (k0 = 0, k1 = w).
// task: (s, I0, J0, k0). 256 threads = 4 waves, each wave a 32x32 quadrant of the tile as
// 2x2 blocks of 16x16 accumulators; K in chunks of 64 staged through LDS (one chunk holds a
// whole step), the next chunk prefetched into registers while the current one is multiplied,
// and the tile's old values loaded at the start. MFMA 16x16x4 operand layout (checked on
// gfx950): lane l gives A[l % 16][l / 16]
// and B[l / 16][l % 16]; accumulator element q of lane l is D[row(l / 16, q)][l % 16] with
// row = 4 (l / 16) + q for the f32 instruction and 4 q + l / 16 for the f64 one.
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ inline acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int l4, int q) { return 4 * q + l4; }
};
template <>
struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ inline acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int l4, int q) { return 4 * l4 + q; }
};

template <typename T>
__global__ void __launch_bounds__(256) k_big_update_mfma(DevSym S, const Task* __restrict__ tasks,
                                                        T* __restrict__ factor, T* __restrict__ cbpool,
                                                        const int32_t* __restrict__ status, int mode) {
    constexpr int KC = 64;                 // a step's whole K (mode 0) in one chunk
    constexpr int PER = KC * TILE / 256;   // entries of A (and of B) per thread and chunk
    typedef typename Mfma<T>::acc_t acc_t;
    __shared__ T As[KC][TILE];
    __shared__ T Bs[KC][TILE];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    if (status[s] != 0) return;
    const int I0 = t.a, J0 = t.b;
    const int k0 = mode == 0 ? t.c : 0;
    const int k1 = mode == 0 ? min(t.c + NB, w) : w;
    const int jlim = mode == 0 ? w : f;
    const int jmin = mode == 0 ? k1 : w;
    T* P = factor + S.panel_off[s];
    T* C = cbpool + (S.cb_off[s] >= 0 ? S.cb_off[s] : 0);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv >> 1, wn = wv & 1;
    const int l16 = lane & 15, l4 = lane >> 4;
    // the destination values, read now so that their latency hides under the products
    T old[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = J0 + wn * 32 + b * 16 + l16;
                const int i = I0 + wm * 32 + a * 16 + Mfma<T>::row(l4, q);
                const bool in = j < jlim && j >= jmin && i < f && i >= j;
                old[a][b][q] = in ? *front_at(P, C, f, w, i, j) : T(0);
            }
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    // this thread's share of a chunk: PER entries of A and of B, (kk, r) = (q / 64, q % 64)
    T pa[PER], pb[PER];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + 256 * u, kk = q >> 6, r = q & 63;
            const int kcol = kc + kk;
            const bool kin = kcol < k1;
            pa[u] = (kin && I0 + r < f) ? P[(int64_t)kcol * f + I0 + r] : T(0);
            pb[u] = (kin && J0 + r < f) ? P[(int64_t)kcol * f + J0 + r] : T(0);
        }
    };
    fetch(k0);
    for (int kc = k0; kc < k1; kc += KC) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + 256 * u, kk = q >> 6, r = q & 63;
            As[kk][r] = pa[u];
            Bs[kk][r] = pb[u];
        }
        __syncthreads();
        if (kc + KC < k1) fetch(kc + KC);
        const int kn = min(KC, k1 - kc);
        for (int k4 = 0; k4 < kn; k4 += 4) {
            const T a0 = As[k4 + l4][wm * 32 + l16], a1 = As[k4 + l4][wm * 32 + 16 + l16];
            const T b0 = Bs[k4 + l4][wn * 32 + l16], b1 = Bs[k4 + l4][wn * 32 + 16 + l16];
            acc[0][0] = Mfma<T>::mac(a0, b0, acc[0][0]);
            acc[0][1] = Mfma<T>::mac(a0, b1, acc[0][1]);
            acc[1][0] = Mfma<T>::mac(a1, b0, acc[1][0]);
            acc[1][1] = Mfma<T>::mac(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int j = J0 + wn * 32 + b * 16 + l16;
            if (j >= jlim || j < jmin) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = I0 + wm * 32 + a * 16 + Mfma<T>::row(l4, q);
                if (i >= f || i < j) continue;
                *front_at(P, C, f, w, i, j) = old[a][b][q] - acc[a][b][q];
            }
        }
}

// per-front statistics of a Cholesky (src/numeric/factorize.jl:13-48) and the smallest
// failed factor column (1-based, 0 = none) into info[0]
__global__ void __launch_bounds__(256) k_chol_stats(DevSym S, const int32_t* __restrict__ status,
                                                   int64_t* __restrict__ stats, int64_t* __restrict__ info) {
    __shared__ int64_t best[256];
    int64_t m = INT64_MAX;
    for (int s = threadIdx.x; s < S.ns; s += blockDim.x) {
        const int64_t fi = status[s];
        int64_t* st = stats + (int64_t)s * NF;
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
    if (threadIdx.x == 0) info[0] = best[0] == INT64_MAX ? 0 : best[0];
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers

template <typename T>
void launch_front_small_chol(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                             const double* nz, T* factor, T* cbpool, int32_t* status, double eps) {
    size_t lds = (size_t)fmax * fmax * sizeof(T);
    hipLaunchKernelGGL(k_front_small_chol<T>, dim3(nfronts), dim3(256), lds, st, S, fronts, nz, factor, cbpool,
                       status, eps);
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
void launch_big_panel(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* DB, int32_t* status,
                      double eps) {
    hipLaunchKernelGGL(k_big_panel<T>, dim3(n), dim3(PANEL_ROWS), 0, st, S, tasks, factor, DB, status, eps);
}

template <typename T>
void launch_big_update(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* cbpool,
                       const int32_t* status, int mode) {
    hipLaunchKernelGGL(k_big_update_mfma<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, cbpool, status, mode);
}

void launch_chol_stats(hipStream_t st, const DevSym& S, const int32_t* status, int64_t* stats, int64_t* info) {
    hipLaunchKernelGGL(k_chol_stats, dim3(1), dim3(256), 0, st, S, status, stats, info);
}

void launch_fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v) {
    if (n <= 0) return;
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks), dim3(256), 0, st, p, n, v);
}

#define INST(T)                                                                                                   \
    template void launch_front_small_chol<T>(hipStream_t, const DevSym&, const int32_t*, int, int, const double*, \
                                             T*, T*, int32_t*, double);                                           \
    template void launch_big_zero_scatter<T>(hipStream_t, const DevSym&, const Task*, int, const double*, T*, T*); \
    template void launch_big_extend_add<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*);                 \
    template void launch_big_panel<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, int32_t*, double);    \
    template void launch_big_update<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, const int32_t*, int);
INST(double)
INST(float)
#undef INST

}  // namespace sds
