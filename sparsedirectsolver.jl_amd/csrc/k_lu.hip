// LU factorization kernels (gfx950) for structure "G": threshold partial pivoting by rows
// inside each front's fully summed block (keep the diagonal when |a_kk| >= u max_i |a_ik|
// over the whole column, else take the largest fully summed row), static perturbation,
// and the Schur update F22 <- F22 - (L21 D) U12.
//
// As the LDL^T kernels, every floating-point operation is the CPU reference's, in the
// same order, with contraction off: in FP64 the row order, D, L, U and the contribution
// blocks are bitwise those of src/reference/lu.jl.
//
// Storage of a front while it is factored: the L panel P (f x w, column-major, ld f,
// the factor buffer) holds the columns < w including the whole fully summed square; the
// rows < w of the columns >= w (U12) are stored transposed in the U panel Q (the ufactor
// buffer, same layout: Q[i * f + j] = F[i, j] for i < w <= j); F22 is the contribution
// block square. At the end P is unit lower (L) and Q holds U^T, unit diagonal.
//
// Reference counterparts: ref_lu! and _lu_front! (src/reference/lu.jl:84-214),
// _lu_choose_row (:216-231), assembly with U^T entries (:114-131), the Schur update
// (:137-145).
#pragma clang fp contract(off)
#include "k_front.h"
#include "sds_common.h"
#include "sds_kernels.h"

namespace sds {

enum { LST_NZERO = 2, LST_NPERT = 3, LST_INFO = 5 };

__device__ inline double lu_jmax(double a, double b) {
    if (a != a || b != b) return a + b;
    return a > b ? a : b;
}

// a front as the LU loop sees it
template <typename T>
struct LuFront {
    T* P;      // columns < w (ld f)
    T* Q;      // rows < w of columns >= w, transposed (ld f); == nullptr: F is the dense f x f P
    int f, w;
    __device__ T& l(int i, int j) const { return P[(int64_t)j * f + i]; }  // j < w
    __device__ T& u(int i, int j) const {                                   // i < w <= j
        return Q ? Q[(int64_t)i * f + j] : P[(int64_t)j * f + i];
    }
    __device__ T& at(int i, int j) const { return j < w ? l(i, j) : u(i, j); }
};

template <int NT>
struct LuRed {
    double v[NT / 64];
    int i[NT / 64];
};

// the LU pivot loop of one front (lu.jl:163-214), all NT threads
template <typename T, int NT>
__device__ void lu_front(const LuFront<T>& Fr, int pivoting, double u, double eps, int* lp, int8_t* kind, T* dd,
                         int* st, LuRed<NT>& rb, double* sh) {
    const int f = Fr.f, w = Fr.w, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int k = 0; k < w; ++k) {
        int r = k;
        if (pivoting) {
            // g = max_{i >= k} |F[i, k]| (NaN-propagating)
            double g = 0.0;
            for (int i = k + tid; i < f; i += NT) g = lu_jmax(g, fabs((double)Fr.l(i, k)));
            for (int o = 32; o > 0; o >>= 1) g = lu_jmax(g, __shfl_xor(g, o));
            if (lane == 0) rb.v[wv] = g;
            __syncthreads();
            g = rb.v[0];
            for (int q = 1; q < NT / 64; ++q) g = lu_jmax(g, rb.v[q]);
            __syncthreads();
            const double akk = fabs((double)Fr.l(k, k));
            if (!(akk >= u * g)) {
                // the largest fully summed row, the first on ties, k if none is larger
                double bv = akk;
                int bi = k;
                for (int i = k + 1 + tid; i < w; i += NT) {
                    const double a = fabs((double)Fr.l(i, k));
                    if (a > bv) {
                        bv = a;
                        bi = i;
                    }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const double ov = __shfl_xor(bv, o);
                    const int oi = __shfl_xor(bi, o);
                    if (ov > bv || (ov == bv && oi < bi)) {
                        bv = ov;
                        bi = oi;
                    }
                }
                if (lane == 0) {
                    rb.v[wv] = bv;
                    rb.i[wv] = bi;
                }
                __syncthreads();
                bv = rb.v[0];
                bi = rb.i[0];
                for (int q = 1; q < NT / 64; ++q)
                    if (rb.v[q] > bv || (rb.v[q] == bv && rb.i[q] < bi)) {
                        bv = rb.v[q];
                        bi = rb.i[q];
                    }
                r = bi;
            }
            __syncthreads();  // every read of the pivot column is done before the swap
        }
        if (r != k) {
            for (int j = tid; j < f; j += NT) {
                T& a = Fr.at(k, j);
                T& b = Fr.at(r, j);
                const T t = a;
                a = b;
                b = t;
            }
            if (tid == 0) {
                const int t = lp[k];
                lp[k] = lp[r];
                lp[r] = t;
            }
            __syncthreads();
        }
        if (tid == 0) {
            double dk = (double)Fr.l(k, k);
            int8_t kd = KIND_1X1;
            if (!(fabs(dk) >= eps) || dk == 0.0) {
                if (dk == 0.0) st[LST_NZERO]++;
                dk = (dk < 0.0 ? -1.0 : 1.0) * eps;
                kd = KIND_PERTURBED;
                st[LST_NPERT]++;
                if (dk == 0.0 && st[LST_INFO] == 0) st[LST_INFO] = k + 1;
            }
            Fr.l(k, k) = T(dk);
            dd[k] = T(dk);
            kind[k] = kd;
            sh[0] = dk;
        }
        __syncthreads();
        const T dk = T(sh[0]);
        // trailing fully summed columns (all rows below k), then U12 (the fully summed rows)
        const int R = f - k - 1, C = w - k - 1;
        for (int q = tid; q < R * C; q += NT) {
            const int j = k + 1 + q / R, i = k + 1 + q % R;
            Fr.l(i, j) = Fr.l(i, j) - Fr.l(i, k) / dk * Fr.l(k, j);
        }
        const int m = f - w;
        for (int q = tid; q < C * m; q += NT) {
            const int j = w + q / C, i = k + 1 + q % C;
            Fr.u(i, j) = Fr.u(i, j) - Fr.l(i, k) / dk * Fr.u(k, j);
        }
        __syncthreads();
        for (int i = k + 1 + tid; i < f; i += NT) Fr.l(i, k) = Fr.l(i, k) / dk;
        for (int j = k + 1 + tid; j < f; j += NT) {
            T& x = Fr.at(k, j);
            x = x / dk;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// small fronts: the dense front in LDS

constexpr int LU_SMALL_W = 128;

template <typename T>
__global__ void __launch_bounds__(256) k_front_small_lu(DevSym S, const int32_t* __restrict__ fronts,
                                                       const double* __restrict__ nz, T* __restrict__ factor,
                                                       T* __restrict__ ufactor, T* __restrict__ cbpool,
                                                       T* __restrict__ dvec, int32_t* __restrict__ piv,
                                                       int8_t* __restrict__ pkind, int64_t* __restrict__ stats,
                                                       int pivoting, double u, double eps, int scaled,
                                                       const double* __restrict__ aux) {
    constexpr int NT = 256;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* F = reinterpret_cast<T*>(smem_raw);
    __shared__ LuRed<NT> rb;
    __shared__ int lp[LU_SMALL_W];
    __shared__ int8_t kind[LU_SMALL_W];
    __shared__ T dd[LU_SMALL_W];
    __shared__ int st[NF];
    __shared__ double sh[2];
    const int s = fronts[blockIdx.x];
    const int f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x;
    if (scaled && aux[0] > 0) eps = eps * aux[0];
    for (int k = tid; k < w; k += NT) {
        lp[k] = k;
        kind[k] = KIND_1X1;
        dd[k] = T(0);
    }
    if (tid < NF) st[tid] = 0;
    assemble_dense<T>(S, s, nz, cbpool, F, f, 1);
    LuFront<T> Fr{F, nullptr, f, w};
    lu_front<T, NT>(Fr, pivoting, u, eps, lp, kind, dd, st, rb, sh);
    // F22 <- F22 - (L21 D) U12, the whole block, each entry's sum in k order
    const int64_t co = S.cb_off[s];
    if (co >= 0) {
        T* C = cbpool + co;
        for (int q = tid; q < m * m; q += NT) {
            const int jj = q / m, ii = q - jj * m;
            T acc = T(0);
            for (int k = 0; k < w; ++k) acc = acc + F[k * f + w + ii] * dd[k] * F[(w + jj) * f + k];
            C[q] = F[(w + jj) * f + w + ii] - acc;
        }
    }
    T* P = factor + S.panel_off[s];
    T* Q = ufactor + S.panel_off[s];
    for (int q = tid; q < f * w; q += NT) {
        const int j = q / f, i = q - j * f;
        P[q] = i > j ? F[q] : (i == j ? T(1) : T(0));
        Q[q] = i > j ? F[i * f + j] : (i == j ? T(1) : T(0));
    }
    for (int k = tid; k < w; k += NT) {
        const int g = c0 + k;
        piv[g] = c0 + lp[k];
        pkind[g] = kind[k];
        dvec[g] = dd[k];
        dvec[S.n + g] = T(0);
    }
    if (tid < NF) stats[(int64_t)s * NF + tid] = st[tid];
}

// ---------------------------------------------------------------------------
// large fronts

// zero a chunk of columns [col0, col0 + ncols) of the front: panel columns of P and Q, or
// contribution-block columns. task: (s, col0, ncols)
template <typename T>
__global__ void __launch_bounds__(256) k_big_lu_zero(DevSym S, const Task* __restrict__ tasks, T* __restrict__ factor,
                                                    T* __restrict__ ufactor, T* __restrict__ cbpool) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], m = f - w;
    T* P = factor + S.panel_off[s];
    T* Q = ufactor + S.panel_off[s];
    T* C = cbpool + (S.cb_off[s] >= 0 ? S.cb_off[s] : 0);
    const int j1 = min(t.a + t.b, f);
    for (int j = t.a; j < j1; ++j) {
        if (j < w) {
            for (int i = threadIdx.x; i < f; i += 256) {
                P[(int64_t)j * f + i] = T(0);
                Q[(int64_t)j * f + i] = T(0);
            }
        } else if (S.cb_off[s] >= 0) {
            for (int i = threadIdx.x; i < m; i += 256) C[(int64_t)(j - w) * m + i] = T(0);
        }
    }
}

// A into a zeroed front, the runs of the assembly entries of panel columns [col0, col0 +
// ncols) (a U^T entry at panel position (i, j) is F[j, i]). task: (s, col0, ncols)
template <typename T>
__global__ void __launch_bounds__(256) k_big_lu_scatter(DevSym S, const Task* __restrict__ tasks,
                                                       const double* __restrict__ nz, T* __restrict__ factor,
                                                       T* __restrict__ ufactor) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    T* P = factor + S.panel_off[s];
    T* Q = ufactor + S.panel_off[s];
    const int c0 = S.super_ptr[s];
    const int j1 = min(t.a + t.b, w);
    for (int j = t.a; j < j1; ++j) {
        const int64_t a0 = S.acol_ptr[c0 + j], a1 = S.acol_ptr[c0 + j + 1];
        for (int64_t k = a0 + threadIdx.x; k < a1; k += 256) {
            const int64_t d = S.asm_dst[k];
            if (k > a0 && S.asm_dst[k - 1] == d) continue;
            double v = 0.0;
            for (int64_t q = k; q < a1 && S.asm_dst[q] == d; ++q) v += nz[S.asm_src[q]];
            if (d >= 0) {
                P[d] += T(v);
            } else {
                const int64_t loc = -(d + 1);
                const int pc = (int)(loc / f), pr = (int)(loc - (int64_t)pc * f);  // F[pc, pr]
                if (pr < w)
                    P[(int64_t)pr * f + pc] += T(v);
                else
                    Q[loc] += T(v);
            }
        }
    }
}

// extend-add of one child per front, the child's whole m x m block. task: (s, child, jj0, njj)
template <typename T>
__global__ void __launch_bounds__(256) k_big_lu_extend_add(DevSym S, const Task* __restrict__ tasks,
                                                          T* __restrict__ factor, T* __restrict__ ufactor,
                                                          T* __restrict__ cbpool) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int c = t.a, mc = S.fnrows[c] - S.fncols[c];
    T* P = factor + S.panel_off[s];
    T* Q = ufactor + S.panel_off[s];
    T* Cs = cbpool + (S.cb_off[s] >= 0 ? S.cb_off[s] : 0);
    const T* Cc = cbpool + S.cb_off[c];
    const int32_t* ri = S.relind + S.relind_ptr[c];
    const int jend = min(t.b + t.c, mc);
    for (int jj = t.b; jj < jend; ++jj) {
        const int pc = ri[jj];
        for (int ii = threadIdx.x; ii < mc; ii += 256) {
            const int pr = ri[ii];
            T* dst = pc < w ? P + (int64_t)pc * f + pr : pr < w ? Q + (int64_t)pr * f + pc
                                                               : Cs + (int64_t)(pc - w) * m + (pr - w);
            *dst = *dst + Cc[(int64_t)jj * mc + ii];
        }
    }
}

// ---------------------------------------------------------------------------
// large fronts, one pivot per pair of launches: k_lu_pivot (one workgroup per front) takes
// the previous pivot's deferred scaling, searches column k, swaps the rows and sets d_k;
// k_lu_update (64 x 64 tiles over every front of the level) applies the rank-1 update of
// step k to the fully summed columns (all rows) and to U12 (the fully summed rows). Each
// entry goes through the same operations, in the same order, as in lu_front above; the
// scaling of column k and row k by d_k moves to the next pivot launch because the update
// tiles read them unscaled.

// the pivot of step k of every front with more than k columns. task: (s)
template <typename T>
__global__ void __launch_bounds__(1024) k_lu_pivot(DevSym S, const Task* __restrict__ tasks, int k,
                                                  T* __restrict__ factor, T* __restrict__ ufactor,
                                                  T* __restrict__ dvec, int32_t* __restrict__ piv,
                                                  int8_t* __restrict__ pkind, int64_t* __restrict__ stats,
                                                  int pivoting, double u, double eps, int scaled,
                                                  const double* __restrict__ aux) {
    constexpr int NT = 1024;
    __shared__ LuRed<NT> rb;
    const int s = tasks[blockIdx.x].s;
    const int f = S.fnrows[s], w = S.fncols[s];
    if (k >= w) return;
    const int c0 = S.super_ptr[s];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (scaled && aux[0] > 0) eps = eps * aux[0];
    LuFront<T> Fr{factor + S.panel_off[s], ufactor + S.panel_off[s], f, w};
    int* lp = piv + c0;
    int8_t* kind = pkind + c0;
    T* dd = dvec + c0;
    int64_t* st = stats + (int64_t)s * NF;
    if (k == 0) {
        for (int q = tid; q < w; q += NT) {
            lp[q] = q;
            kind[q] = KIND_1X1;
            dd[q] = T(0);
            dvec[S.n + c0 + q] = T(0);
        }
        if (tid < NF) st[tid] = 0;
    }
    // The usual step in one pass: column k's entries k..f-1 in registers, loaded together with
    // step k - 1's column and row (which are divided by its pivot here) and reduced once for the
    // threshold test's column maximum and the row search; the general loops below serve fronts
    // with more than CR x NT rows.
    constexpr int CR = 8;
    __shared__ double sg[NT / 64];
    int r = k;
    if (f - k <= CR * NT) {
        T x[CR], cv[CR], rv[CR];
        const T* colk = &Fr.l(0, k);
#pragma unroll
        for (int q = 0; q < CR; ++q) x[q] = colk[min(k + tid + NT * q, f - 1)];
        const double akk = fabs((double)colk[k]);
        if (k > 0) {
            const T dp = dd[k - 1];
#pragma unroll
            for (int q = 0; q < CR; ++q) {
                const int i = min(k + tid + NT * q, f - 1);
                cv[q] = Fr.l(i, k - 1);
                rv[q] = Fr.at(k - 1, i);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < CR; ++q) {
                const int i = k + tid + NT * q;
                if (i < f) {
                    Fr.l(i, k - 1) = cv[q] / dp;
                    Fr.at(k - 1, i) = rv[q] / dp;
                }
            }
        }
        if (pivoting) {
            double g = 0.0, bv = akk;
            int bi = k;
#pragma unroll
            for (int q = 0; q < CR; ++q) {
                const int i = k + tid + NT * q;
                if (i < f) {
                    const double a = fabs((double)x[q]);
                    g = lu_jmax(g, a);
                    if (i > k && i < w && a > bv) {
                        bv = a;
                        bi = i;
                    }
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                g = lu_jmax(g, __shfl_xor(g, o));
                const double ov = __shfl_xor(bv, o);
                const int oi = __shfl_xor(bi, o);
                if (ov > bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (lane == 0) {
                sg[wv] = g;
                rb.v[wv] = bv;
                rb.i[wv] = bi;
            }
            __syncthreads();
            g = sg[0];
            bv = rb.v[0];
            bi = rb.i[0];
            for (int q = 1; q < NT / 64; ++q) {
                g = lu_jmax(g, sg[q]);
                if (rb.v[q] > bv || (rb.v[q] == bv && rb.i[q] < bi)) {
                    bv = rb.v[q];
                    bi = rb.i[q];
                }
            }
            if (!(akk >= u * g)) r = bi;
        }
        __syncthreads();
    } else {
        if (k > 0) {
            // step k - 1's column and row, divided by its pivot
            const T dp = dd[k - 1];
            for (int i = k + tid; i < f; i += NT) Fr.l(i, k - 1) = Fr.l(i, k - 1) / dp;
            for (int j = k + tid; j < f; j += NT) {
                T& x = Fr.at(k - 1, j);
                x = x / dp;
            }
        }
        __syncthreads();
        if (pivoting) {
            double g = 0.0;
            for (int i = k + tid; i < f; i += NT) g = lu_jmax(g, fabs((double)Fr.l(i, k)));
            for (int o = 32; o > 0; o >>= 1) g = lu_jmax(g, __shfl_xor(g, o));
            if (lane == 0) rb.v[wv] = g;
            __syncthreads();
            g = rb.v[0];
            for (int q = 1; q < NT / 64; ++q) g = lu_jmax(g, rb.v[q]);
            __syncthreads();
            const double akk = fabs((double)Fr.l(k, k));
            if (!(akk >= u * g)) {
                double bv = akk;
                int bi = k;
                for (int i = k + 1 + tid; i < w; i += NT) {
                    const double a = fabs((double)Fr.l(i, k));
                    if (a > bv) {
                        bv = a;
                        bi = i;
                    }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const double ov = __shfl_xor(bv, o);
                    const int oi = __shfl_xor(bi, o);
                    if (ov > bv || (ov == bv && oi < bi)) {
                        bv = ov;
                        bi = oi;
                    }
                }
                if (lane == 0) {
                    rb.v[wv] = bv;
                    rb.i[wv] = bi;
                }
                __syncthreads();
                bv = rb.v[0];
                bi = rb.i[0];
                for (int q = 1; q < NT / 64; ++q)
                    if (rb.v[q] > bv || (rb.v[q] == bv && rb.i[q] < bi)) {
                        bv = rb.v[q];
                        bi = rb.i[q];
                    }
                r = bi;
            }
            __syncthreads();
        }
    }
    if (r != k) {
        for (int j = tid; j < f; j += NT) {
            T& a = Fr.at(k, j);
            T& b = Fr.at(r, j);
            const T t = a;
            a = b;
            b = t;
        }
        if (tid == 0) {
            const int t = lp[k];
            lp[k] = lp[r];
            lp[r] = t;
        }
        __syncthreads();
    }
    if (tid == 0) {
        double dk = (double)Fr.l(k, k);
        int8_t kd = KIND_1X1;
        if (!(fabs(dk) >= eps) || dk == 0.0) {
            if (dk == 0.0) st[LST_NZERO]++;
            dk = (dk < 0.0 ? -1.0 : 1.0) * eps;
            kd = KIND_PERTURBED;
            st[LST_NPERT]++;
            if (dk == 0.0 && st[LST_INFO] == 0) st[LST_INFO] = k + 1;
        }
        Fr.l(k, k) = T(dk);
        dd[k] = T(dk);
        kind[k] = kd;
    }
}

// the rank-1 update of step k: tile (s, I0, J0) of 64 x 64 entries, J0 < w: fully summed
// columns (rows I0.. of every row > k), J0 >= w: U12 (rows I0.. < w, stored transposed)
template <typename T>
__global__ void __launch_bounds__(256) k_lu_update(DevSym S, const Task* __restrict__ tasks, int k,
                                                   T* __restrict__ factor, T* __restrict__ ufactor,
                                                   const T* __restrict__ dvec) {
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s];
    if (k >= w) return;
    const int I0 = t.a, J0 = t.b;
    LuFront<T> Fr{factor + S.panel_off[s], ufactor + S.panel_off[s], f, w};
    const T dk = dvec[S.super_ptr[s] + k];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    if (J0 < w) {
        // lanes along the rows: column-major panel
        const int i = I0 + tx;
        if (i <= k || i >= f) return;
        // the thread's sixteen columns: loads in flight together (clamped), then the stores;
        // row k is not written here
        const int jb = max(J0, k + 1) + ty, j1 = min(J0 + 64, w);
        if (jb >= j1) return;
        const T lk = Fr.l(i, k) / dk;
        T a[16], b[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int j = min(jb + 4 * q, j1 - 1);
            a[q] = Fr.l(i, j);
            b[q] = Fr.l(k, j);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (jb + 4 * q < j1) Fr.l(i, jb + 4 * q) = a[q] - lk * b[q];
    } else {
        // lanes along the columns: U12 is stored transposed; the thread's sixteen rows as above
        const int j = J0 + tx;
        if (j >= f) return;
        const int ib = max(I0, k + 1) + ty, i1 = min(I0 + 64, w);
        if (ib >= i1) return;
        const T rk = Fr.u(k, j);
        T a[16], l[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int i = min(ib + 4 * q, i1 - 1);
            l[q] = Fr.l(i, k);
            a[q] = Fr.u(i, j);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (ib + 4 * q < i1) Fr.u(ib + 4 * q, j) = a[q] - (l[q] / dk) * rk;
    }
}

// after the last pivot of a front: its deferred scaling, then U^T (the upper triangle of the
// fully summed square moves to Q) and the unit diagonals. task: (s)
template <typename T>
__global__ void __launch_bounds__(1024) k_lu_finish(DevSym S, const Task* __restrict__ tasks,
                                                   T* __restrict__ factor, T* __restrict__ ufactor,
                                                   const T* __restrict__ dvec, int32_t* __restrict__ piv) {
    constexpr int NT = 1024;
    const int s = tasks[blockIdx.x].s;
    const int f = S.fnrows[s], w = S.fncols[s], c0 = S.super_ptr[s], tid = threadIdx.x;
    T* P = factor + S.panel_off[s];
    T* Q = ufactor + S.panel_off[s];
    LuFront<T> Fr{P, Q, f, w};
    const T dp = dvec[c0 + w - 1];
    for (int i = w + tid; i < f; i += NT) Fr.l(i, w - 1) = Fr.l(i, w - 1) / dp;
    for (int j = w + tid; j < f; j += NT) {
        T& x = Fr.at(w - 1, j);
        x = x / dp;
    }
    __syncthreads();
    // U^T: eight entries per thread per pass, their loads in flight together (clamped to the
    // square), then the stores of those below the diagonal
    const int ww = w * w;
    for (int q0 = 0; q0 < ww; q0 += 8 * NT) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = min(q0 + tid + u * NT, ww - 1), j = q / w, i = q - j * w;
            v[u] = P[(int64_t)i * f + j];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + tid + u * NT, j = q / w, i = q - j * w;
            if (q < ww && i > j) Q[(int64_t)j * f + i] = v[u];
        }
    }
    __syncthreads();
    for (int q = tid; q < w * w; q += NT) {
        const int j = q / w, i = q - j * w;
        if (i <= j) {
            P[(int64_t)j * f + i] = i == j ? T(1) : T(0);
            Q[(int64_t)j * f + i] = i == j ? T(1) : T(0);
        }
    }
    for (int q = tid; q < w; q += NT) piv[c0 + q] = c0 + piv[c0 + q];
}

// F22 <- F22 - (L21 D) U12 by 64 x 64 tiles (I0, J0 front rows/columns >= w), each entry's
// sum over k = 0, 1, ..., w - 1 in order. L21 D and U12 in chunks of 16 columns in LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_lu_cb(DevSym S, const Task* __restrict__ tasks,
                                              const T* __restrict__ factor, const T* __restrict__ ufactor,
                                              T* __restrict__ cbpool, const T* __restrict__ dvec) {
    constexpr int KC = 16;
    __shared__ T Ws[KC][TILE];
    __shared__ T Us[KC][TILE];
    const Task t = tasks[blockIdx.x];
    const int s = t.s, f = S.fnrows[s], w = S.fncols[s], m = f - w;
    const int I0 = t.a, J0 = t.b;
    const int c0 = S.super_ptr[s];
    const T* L = factor + S.panel_off[s];
    const T* Q = ufactor + S.panel_off[s];
    T* C = cbpool + S.cb_off[s];
    const T* dd = dvec + c0;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
    // the thread's four (kk, c) staging slots, kk = (tid >> 6) + 4 u, c = tid & 63: the raw
    // values of chunk k0 + KC are loaded, clamped, while chunk k0 is multiplied
    const int c = tid & 63, ic = min(I0 + c, f - 1), jc = min(J0 + c, f - 1);
    T lv[4], dv[4], uv0[4];
    auto load_raw = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kc = min(k0 + (tid >> 6) + 4 * u, w - 1);
            lv[u] = L[(int64_t)kc * f + ic];
            dv[u] = dd[kc];
            uv0[u] = Q[(int64_t)kc * f + jc];
        }
    };
    load_raw(0);
    for (int k0 = 0; k0 < w; k0 += KC) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = (tid >> 6) + 4 * u, k = k0 + kk;
            Ws[kk][c] = (k < w && I0 + c < f) ? lv[u] * dv[u] : T(0);
            Us[kk][c] = (k < w && J0 + c < f) ? uv0[u] : T(0);
        }
        __syncthreads();
        if (k0 + KC < w) load_raw(k0 + KC);
        const int kn = min(KC, w - k0);
        for (int kk = 0; kk < kn; ++kk) {
            T wv[4], uv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) wv[a] = Ws[kk][4 * ty + a];
#pragma unroll
            for (int b = 0; b < 4; ++b) uv[b] = Us[kk][4 * tx + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = acc[a][b] + wv[a] * uv[b];
        }
        __syncthreads();
    }
    // the tile's old values loaded together (clamped), then the masked stores
    T old[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = min(I0 + 4 * ty + a, f - 1), j = min(J0 + 4 * tx + b, f - 1);
            old[a][b] = C[(int64_t)(j - w) * m + (i - w)];
        }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = I0 + 4 * ty + a, j = J0 + 4 * tx + b;
            if (i < f && j < f) C[(int64_t)(j - w) * m + (i - w)] = old[a][b] - acc[a][b];
        }
}

// ---------------------------------------------------------------------------

template <typename T>
void launch_front_small_lu(hipStream_t st, const DevSym& S, const int32_t* fronts, int nfronts, int fmax,
                           const double* nz, T* factor, T* ufactor, T* cbpool, T* dvec, int32_t* piv, int8_t* kind,
                           int64_t* stats, int pivoting, double u, double eps, int scaled, const double* aux) {
    size_t lds = (size_t)fmax * fmax * sizeof(T);
    hipLaunchKernelGGL(k_front_small_lu<T>, dim3(nfronts), dim3(256), lds, st, S, fronts, nz, factor, ufactor, cbpool,
                       dvec, piv, kind, stats, pivoting, u, eps, scaled, aux);
}

template <typename T>
void launch_big_lu_zero(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* ufactor, T* cbpool) {
    hipLaunchKernelGGL(k_big_lu_zero<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, ufactor, cbpool);
}

template <typename T>
void launch_big_lu_scatter(hipStream_t st, const DevSym& S, const Task* tasks, int n, const double* nz, T* factor,
                           T* ufactor) {
    hipLaunchKernelGGL(k_big_lu_scatter<T>, dim3(n), dim3(256), 0, st, S, tasks, nz, factor, ufactor);
}

template <typename T>
void launch_big_lu_extend_add(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* ufactor,
                              T* cbpool) {
    hipLaunchKernelGGL(k_big_lu_extend_add<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, ufactor, cbpool);
}

template <typename T>
void launch_lu_pivot(hipStream_t st, const DevSym& S, const Task* tasks, int n, int k, T* factor, T* ufactor,
                     T* dvec, int32_t* piv, int8_t* kind, int64_t* stats, int pivoting, double u, double eps,
                     int scaled, const double* aux) {
    hipLaunchKernelGGL(k_lu_pivot<T>, dim3(n), dim3(1024), 0, st, S, tasks, k, factor, ufactor, dvec, piv, kind,
                       stats, pivoting, u, eps, scaled, aux);
}

template <typename T>
void launch_lu_update(hipStream_t st, const DevSym& S, const Task* tasks, int n, int k, T* factor, T* ufactor,
                      const T* dvec) {
    hipLaunchKernelGGL(k_lu_update<T>, dim3(n), dim3(256), 0, st, S, tasks, k, factor, ufactor, dvec);
}

template <typename T>
void launch_lu_finish(hipStream_t st, const DevSym& S, const Task* tasks, int n, T* factor, T* ufactor,
                      const T* dvec, int32_t* piv) {
    hipLaunchKernelGGL(k_lu_finish<T>, dim3(n), dim3(1024), 0, st, S, tasks, factor, ufactor, dvec, piv);
}

template <typename T>
void launch_lu_cb(hipStream_t st, const DevSym& S, const Task* tasks, int n, const T* factor, const T* ufactor,
                  T* cbpool, const T* dvec) {
    hipLaunchKernelGGL(k_lu_cb<T>, dim3(n), dim3(256), 0, st, S, tasks, factor, ufactor, cbpool, dvec);
}

#define INST(T)                                                                                                      \
    template void launch_front_small_lu<T>(hipStream_t, const DevSym&, const int32_t*, int, int, const double*, T*, \
                                           T*, T*, T*, int32_t*, int8_t*, int64_t*, int, double, double, int,       \
                                           const double*);                                                          \
    template void launch_big_lu_zero<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, T*);                  \
    template void launch_big_lu_scatter<T>(hipStream_t, const DevSym&, const Task*, int, const double*, T*, T*);    \
    template void launch_big_lu_extend_add<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, T*);            \
    template void launch_lu_cb<T>(hipStream_t, const DevSym&, const Task*, int, const T*, const T*, T*, const T*); \
    template void launch_lu_pivot<T>(hipStream_t, const DevSym&, const Task*, int, int, T*, T*, T*, int32_t*,      \
                                     int8_t*, int64_t*, int, double, double, int, const double*);                   \
    template void launch_lu_update<T>(hipStream_t, const DevSym&, const Task*, int, int, T*, T*, const T*);         \
    template void launch_lu_finish<T>(hipStream_t, const DevSym&, const Task*, int, T*, T*, const T*, int32_t*);
INST(double)
INST(float)
#undef INST

}  // namespace sds
